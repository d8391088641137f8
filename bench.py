"""Benchmark: ms per Newton iteration of the Stiff-GIPC hot path on one MI355X.

Default workload (the stiff/soft hybrid impact): 512 affine boxes of 6^3 hex
cells (1,296 tets each, kappa 1e11) fired at 10 m/s into a soft Neo-Hookean
bed of 64x64x40 hex cells (983,040 tets, E = 5e4 Pa, bottom layer fixed);
1,646,592 tets, dt = 0.005 s, dhat = 5e-4 m, friction mu = 0.3.

A bench "step" is one Newton iteration that solves (assembly, MAS build,
PCG, CCD bounds, line search) of the running time integration.  The scene is
integrated for --settle time steps (untimed, every Newton loop run to its
stop test) so that the timed iterations are mid-impact.  Each solver mode
then runs --warmup untimed and --steps timed Newton iterations from the same
state:

  fp64_ref       the reference's plain FP64 solve: FP64 SpMV/PCG, FP32 MAS
                 apply, stop on the preconditioned norm r.z every 10
  fp64_true      the same solve stopped on the FP64 true relative residual
                 ||b - A x|| / ||b|| <= 1e-4 (the headline's baseline)
  mixed_true     FP32 matrix, vectors and MAS apply in the inner PCG, FP64
                 residual refinement, same true-residual stop
  mixed_learned  mixed_true with the network-corrected coarse warm start
                 (the headline's method; `value`)

`value` is mixed_learned's ms per Newton iteration; `ratio_vs_fp64` divides it
by fp64_true's, both at the same FP64 true residual 1e-4.

Prints ONE JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "stiff-gipc_amd"))

METRIC = "ms/Newton step + PCG iters at 1e-4 res: plain FP64 MAS vs FP32+learned"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def upper_blocks(sz):
    """Unique upper-triangle 3x3 blocks of the assembled system: the static
    pattern is stored with both triangles (nnzb = 2 * upper - n_nodes) and
    every distinct contact block in both orientations."""
    return (sz["nnzb"] + sz["n_nodes"]) // 2 + sz["coo_blocks"] // 2


def algorithmic_bytes(name, sz, nparts, fp32):
    """Bytes one launch must move at minimum, each operand read or written
    once, with the matrix counted as its unique upper-triangle blocks.

    spmv (the PCG product, which also forms p = z + beta p_prev): per upper
    block its values (72 B FP64 / 36 B FP32) and a row and a column index
    (8 B); per node z and p_prev read, p and A p written (4 vectors of 24 B
    FP64 / 12 B FP32).
    mas_apply_l0: the level-0 bank inverses as the reference stores them
    (upper 3x3 blocks of the symmetric 48x48 in FP32: 136 blocks x 36 B =
    4,896 B per bank); per DOF x, p, A p, r read, x, r written (FP64 or
    FP32) and the FP32 restriction written."""
    nv = sz["n_nodes"]
    w = 4 if fp32 else 8
    if name == "spmv":
        return upper_blocks(sz) * (9 * w + 8) + nv * 4 * 3 * w
    if name == "mas_apply_l0":
        return nparts * 4896 + 3 * nv * (6 * w + 4)
    return 0


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (calibrated against a
    plain copy); None when absent."""
    path = os.path.join(ROOT, "profiles", f"spmv_{workload}_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    parts = line.split(":", 1)[1].split()
                    if len(parts) == 2 and parts[0] == kernel:
                        return int(parts[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def build_scene(args, S):
    if args.workload == "hybrid":
        sc = S.hybrid_scene()
        desc = (f"hybrid impact: 512 affine boxes (6^3 cells, 1,296 tets each, kappa 1e11) at "
                f"10 m/s into a 64x64x40-cell bed (983,040 tets, E 5e4 Pa, bottom layer fixed); "
                f"{sc.n_tets} tets, {sc.n_verts} verts; dt 0.005, dhat 5e-4, mu 0.3")
        return sc, 1, desc
    if args.workload == "cloth_abd":
        sc = S.cloth_scene(n=256, size=1.0, height=0.3, n_boxes=16, box_size=0.1, box_cells=3,
                           E=1e6, bend_stiffness=1e-4, thickness=1e-3, abd_boxes=True,
                           abd_kappa=1e11)
        desc = (f"cloth sheet 256x256 verts ({sc.tris.shape[0]} triangles, 1 m), membrane "
                f"E 1e6 Pa, bending 1e-4, thickness 1e-3, dropped 0.3 m over 16 affine boxes "
                f"(kappa 1e11); dt 0.01, dhat 1e-3")
        return sc, 60, desc
    sc = S.cubes_scene(n_side=args.n_side, cells=args.cells, size=0.2, spacing=0.6, fall=0.5,
                       seed=0)
    desc = (f"{args.n_side ** 3} tet cubes of {args.cells}^3 hex cells ({sc.n_tets} tets, "
            f"{sc.n_verts} verts), E log-uniform [1e5,1e9] Pa, nu 0.4, 0.5 m drop onto ground, "
            f"dt 0.01, dhat 1e-3")
    return sc, 40, desc


def cpu_baseline(sc, state, iters, threads):
    """The oracle (FP64 C restatement, OpenMP over `threads` host threads)
    from the same settled state: ms per Newton iteration."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    O.set_threads(threads)
    w = O.World(sc)
    w.set_state(*state)
    c0 = time.perf_counter()
    w.advance(iters)
    c1 = time.perf_counter()
    return (c1 - c0) * 1e3 / iters, O.get_threads()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed Newton iterations per mode")
    ap.add_argument("--warmup", type=int, default=3, help="untimed Newton iterations per mode")
    ap.add_argument("--workload", default="hybrid", choices=["hybrid", "cubes64", "cloth_abd"],
                    help="hybrid: 512 affine boxes fired into a 983k-tet bed (default); cubes64: "
                         "64 tet cubes; cloth_abd: 256x256 cloth over 16 affine boxes")
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed time steps before the timed modes (default 1 / 40 / 60)")
    ap.add_argument("--n-side", type=int, default=4)
    ap.add_argument("--cells", type=int, default=8)
    ap.add_argument("--mas-fp64", action="store_true", help="all-FP64 MAS instead of FP32 apply")
    ap.add_argument("--morton", action="store_true",
                    help="spatially re-sort the scene (Morton vertex order, centroid-key tet "
                         "order, MAS banks of 16 consecutive vertices) before the run")
    ap.add_argument("--cpu-iters", type=int, default=1, help="oracle Newton iterations timed")
    ap.add_argument("--cpu-threads", type=int, default=0, help="oracle threads (0: all cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="headline", choices=["headline", "all", "ref"],
                    help="headline: fp64_ref, fp64_true, mixed_true, mixed_learned; all: also "
                         "the FP64 solves from the affine and the learned guess; ref: fp64_ref")
    ap.add_argument("--train-iters", type=int, default=12,
                    help="Newton solves whose coarse systems train the network")
    ap.add_argument("--train-epochs", type=int, default=400)
    ap.add_argument("--train-crop", default=None,
                    help="cubes workload: collect the training solves on a crop of the scene, "
                         "NA,NB,NC cubes along x, y (vertical), z (e.g. 4,2,4: 32 cubes), "
                         "settled for the same number of steps, instead of the scene itself")
    args = ap.parse_args()

    if args.train_crop and args.workload != "cubes64":
        print("bench.py: --train-crop applies to the cubes64 workload", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is design-only", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sgipc
    from sgipc import scene as S

    log("device:", sgipc.device_info())
    sc, settle_default, workload = build_scene(args, S)
    if args.settle is None:
        args.settle = settle_default
    if args.morton:
        t0 = time.time()
        sc = sgipc.spatial_sort(sc).scene
        workload += "; Morton re-sorted (banks of 16 consecutive vertices)"
        log(f"spatial sort: {time.time() - t0:.2f} s")
    sc.mas_precision = sgipc.SGIPC_MAS_FP64 if args.mas_fp64 else sgipc.SGIPC_MAS_FP32
    sc.pcg_rel_tol = 1e-4
    nparts = int(sc.part_id.max()) + 1
    log(f"scene: {sc.n_verts} verts, {sc.n_tets} tets, {sc.faces.shape[0]} faces, "
        f"{sc.edges.shape[0]} edges, {nparts} banks, newton cap {sc.newton_max_iter}")
    t0 = time.time()
    sim = sgipc.Simulation(sc)
    log(f"create: {time.time() - t0:.2f} s")

    # ---- settle (untimed): whole time steps, each Newton loop to its stop test
    t0 = time.time()
    settle = []
    for k in range(args.settle):
        st = sim.time_step()
        x = sim.positions()
        settle.append({"newton": st["newton_iterations"], "pcg": st["pcg_iterations"],
                       "capped": st["capped"]})
        log(f"settle step {k + 1}/{args.settle}: newton={st['newton_iterations']} "
            f"pcg={st['pcg_iterations']} capped={st['capped']} ymin={x[:, 1].min():.4f} "
            f"({time.time() - t0:.1f} s)")
    state = (sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)

    # ---- training data for the learned coarse correction (untimed): solves
    # of the settled scene with the affine warm start at a tight tolerance, so
    # the targets carry the coupling between aggregates; full-batch Adam on GPU
    train = None
    if args.modes != "ref" and args.train_iters > 0:
        t0 = time.time()
        tsim = sim
        if args.train_crop:
            crop = tuple(int(v) for v in args.train_crop.split(","))
            csc = S.cubes_scene(n_side=args.n_side, cells=args.cells, size=0.2, spacing=0.6,
                                fall=0.5, seed=0, crop=crop)
            csc.mas_precision = sc.mas_precision
            tsim = sgipc.Simulation(csc)
            for _ in range(args.settle):
                tsim.time_step()
            log(f"training crop {crop}: {csc.n_verts} verts, {csc.n_tets} tets, settled "
                f"{args.settle} steps ({time.time() - t0:.1f} s)")
        tsim.set_pcg(rel_tol=1e-8, use_rz=True, solver_precision=sgipc.SGIPC_SOLVE_FP64)
        tsim.learned_mode(1)
        tsim.warm_start_adaptive(False)
        F, T = [], []
        for _ in range(args.train_iters):
            tsim.advance(1)
            f, t = tsim.learned_samples()
            F.append(f)
            T.append(t)
        F, T = np.concatenate(F), np.concatenate(T)
        t1 = time.time()
        loss = tsim.mlp_train(F, T, epochs=args.train_epochs, lr=1e-3)
        t2 = time.time()
        if tsim is not sim:
            sim.mlp_weights(tsim.mlp_weights())
            tsim.close()
        sim.learned_mode(0)
        sim.warm_start_adaptive(True)
        train = {"samples": int(F.shape[0]), "collect_s": round(t1 - t0, 2),
                 "train_s": round(t2 - t1, 2), "epochs": args.train_epochs,
                 "loss_untrained": float((T.astype(np.float64) ** 2).mean()),
                 "loss_trained": loss}
        if args.train_crop:
            train["crop"] = args.train_crop
        log(f"learned correction: {train}")

    ref = dict(rel_tol=1e-4, use_rz=True, solver_precision=sgipc.SGIPC_SOLVE_FP64)
    true64 = dict(rel_tol=1e-4, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_FP64)
    true32 = dict(rel_tol=1e-4, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_MIXED)
    modes = [("fp64_ref", ref, 0), ("fp64_true", true64, 0), ("mixed_true", true32, 0),
             ("mixed_learned", true32, 2)]
    if args.modes == "all":
        modes += [("fp64_affine", ref, 1), ("fp64_learned", ref, 2)]
    if args.modes == "ref":
        modes = modes[:1]
    value_mode = "mixed_learned" if train is not None else "fp64_ref"
    base_mode = "fp64_true" if train is not None else "fp64_ref"

    results, kern, advs = {}, {}, {}
    for name, cfg, lmode in modes:
        if lmode == 2 and train is None:
            continue
        sim.set_state(*state)
        sim.set_pcg(**cfg)
        sim.learned_mode(lmode)
        if args.warmup:
            sim.advance(args.warmup)
        # HIP events around every 10th launch of the profiled kernels, in every
        # mode's timed region alike (an event pair adds microseconds of stream time)
        sim.profile(10)
        t0 = time.perf_counter()
        adv = sim.advance(args.steps)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        n = args.steps
        r = {"ms_per_newton": round(ms, 4),
             "pcg_iterations_per_newton": round(adv["pcg_iterations"] / n, 2),
             "max_rel_residual": adv["max_pcg_rel_residual"],
             "fp64_fallbacks": adv["solver_fallbacks"],
             "capped_steps": adv["capped_steps"],
             "time_steps_begun": adv["time_steps_begun"],
             "phases_ms": {k: round(adv["ms_" + k] / n, 3)
                           for k in ("assembly", "solve", "ccd", "line_search")},
             "line_search_steps": adv["line_search_steps"],
             "ccd_limited_iterations": adv["ccd_limited"]}
        if lmode:
            r["warm_starts"] = adv["warm_starts"]
            r["warm_fallbacks"] = adv["warm_fallbacks"]
        results[name] = r
        advs[name] = (adv, ms)
        log(f"timed [{name}]: {n} Newton iterations in {(t1 - t0):.3f} s -> {ms:.3f} ms/iter; "
            f"pcg/iter={adv['pcg_iterations'] / n:.1f} "
            f"max_rel_res={adv['max_pcg_rel_residual']:.2e} phases={r['phases_ms']} "
            f"steps_begun={adv['time_steps_begun']} capped={adv['capped_steps']} "
            f"fallbacks={adv['solver_fallbacks']}")
        kern[name] = {}
        for kname in ("spmv", "mas_apply_l0", "fem", "mas_invert"):
            nl, tot = sim.profile_query(kname)
            if nl:
                kern[name][kname] = (nl, tot)
                log(f"  kernel {kname}: {nl} launches, {tot:.3f} ms total, "
                    f"{tot / nl * 1e3:.2f} us avg")
        sim.profile(0)
    sim.learned_mode(0)
    sz = sim.sizes()
    log(f"sizes={sz}")

    adv, ms = advs[value_mode]
    base_ms = advs[base_mode][1]

    # ---- roofline of the dominant profiled kernel of the value mode
    fp32 = value_mode.startswith("mixed")
    hbm = {k: v for k, v in kern[value_mode].items() if k in ("spmv", "mas_apply_l0")}
    roof = None
    if hbm:
        dom = max(hbm, key=lambda k: hbm[k][1])
        nl, tot = hbm[dom]
        avg_s = tot / nl * 1e-3
        byt = algorithmic_bytes(dom, sz, nparts, fp32)
        ach = byt / avg_s / 1e9
        default_shape = (args.settle == settle_default and args.n_side == 4 and args.cells == 8
                         and not args.morton)
        traffic, src = measured_traffic(dom + ("_fp32" if fp32 else ""), args.workload) \
            if default_shape else (None, None)
        roof = {"bound": "hbm", "kernel": dom + (" (FP32 inner PCG)" if fp32 else ""),
                "achieved": round(ach, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": src, "algorithmic_bytes_per_launch": int(byt),
                "upper_blocks": int(upper_blocks(sz)), "avg_launch_us": round(avg_s * 1e6, 3),
                "launches_timed": nl}

    # ---- CPU baseline: the oracle from the same settled state, all host cores
    cpu = None
    if not args.no_cpu and args.cpu_iters > 0:
        try:
            # the host cores this process may use (OMP_NUM_THREADS when set)
            threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS") or 0) \
                or (os.cpu_count() or 1)
            cms, used = cpu_baseline(sc, state, args.cpu_iters, threads)
            cpu = {"value": round(cms, 3), "unit": "ms/Newton step", "cores": used,
                   "kind": "port",
                   "sample": f"{args.cpu_iters} Newton iteration(s) of the same workload from the "
                             f"settled state (oracle, FP64 throughout, plain FP64 MAS-PCG with "
                             f"the r.z stop), OpenMP over {used} host threads"}
            log(f"cpu baseline: {cms:.1f} ms/iter on {used} threads")
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e!r}")

    out = {
        "metric": METRIC,
        "value": round(ms, 4),
        "unit": "ms/Newton step",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if fp32 else "f64",
        "data": "synthetic",
        "value_mode": value_mode,
        "baseline_mode": base_mode,
        "fp64_ms_per_newton": round(base_ms, 4),
        "ratio_vs_fp64": round(ms / base_ms, 4),
        "pcg_iterations_per_newton": round(adv["pcg_iterations"] / args.steps, 2),
        "max_pcg_rel_residual": adv["max_pcg_rel_residual"],
        "modes": results,
        "settle_steps": settle,
        "learned_training": train,
        "config": {
            "workload": workload + f", {args.settle} settle step(s)",
            "mode": ("FP32 SpMV/MAS/PCG vectors with FP64 residual refinement and the learned "
                     "coarse warm start" if fp32 else "FP64 PCG/SpMV, FP32 MAS apply"),
            "pcg_rel_tol": 1e-4,
            "residual": "FP64 true relative residual" if fp32 else "r.z (reference test)",
            "newton_cap": sc.newton_max_iter,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
