"""Benchmark: ms per Newton iteration of the Stiff-GIPC hot path on one MI355X.

Workload (the metric's mid configuration): 64 tet cubes of 8^3 hex cells
(3072 tets each, 196,608 tets, 46,656 vertices) on a 4x4x4 grid with seeded
+/-15% jitter and random rotations, Young's modulus log-uniform in
[1e5, 1e9] Pa, nu = 0.4, dropped 0.5 m onto the ground; dt = 0.01 s,
dhat = 1e-3 m, MAS-PCG to 1e-4 relative (preconditioned) residual.

A bench "step" is one Newton iteration that solves (assembly, MAS build,
PCG, CCD bounds, line search) of the running time integration.  The scene is
first integrated for --settle time steps (untimed) so the timed iterations
include ground and cube-cube contacts.  Value: FP64 assembly/SpMV/PCG with the
reference's FP32 MAS apply (the reference's "plain FP64" path); the "modes"
object adds the FP64 true-residual and mixed-precision solves and the solves
continued from the affine coarse guess with and without the learned
correction (trained, untimed, on the settled scene's own Newton systems).

Prints ONE JSON line on rank 0.
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


def algorithmic_bytes(name, sz, nparts):
    """Bytes one launch must move at minimum (each operand read or written once)."""
    nv, nt, nnzb, ncoo = sz["n_nodes"], sz["n_tets"], sz["nnzb"], sz["coo_blocks"]
    if name == "spmv":
        # the PCG SpMV forms p = z + beta p_prev on the fly and computes
        # y = A p: static BSR 72 B block + 4 B column and its row pointers;
        # contact rows (row-sorted) 72 B block + 4 B column and row pointers;
        # per node z and p_prev read, p and y written (24 B each)
        return nnzb * 76 + ncoo * 76 + 2 * (nv + 1) * 4 + nv * 96
    if name == "mas_apply_l0":
        # FP32 48x48 inverse per level-0 bank; per dof: read x, p, Ap, r, write
        # x, r (FP64) and Z (FP32); slot map
        return nparts * (48 * 48 * 4 + 16 * 4) + 3 * nv * (6 * 8 + 4)
    if name == "fem":
        # tet ids, Dm^-1, vol/mu/lambda, 12 coordinates, 16 slot ids,
        # 16 blocks accumulated, 12 gradient entries accumulated
        return nt * (16 + 72 + 24 + 96 + 64 + 16 * 72 + 12 * 8)
    return 0


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on the default workload, from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary (calibrated against a
    plain copy); None when absent."""
    path = os.path.join(ROOT, "profiles", "spmv_64cubes_hbm_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    parts = line.split(":", 1)[1].split()
                    if len(parts) == 2 and parts[0] == kernel:
                        return int(parts[1])
    except OSError:
        pass
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed Newton iterations")
    ap.add_argument("--warmup", type=int, default=3, help="untimed Newton iterations")
    ap.add_argument("--workload", default="cubes64", choices=["cubes64", "hybrid", "cloth_abd"],
                    help="cubes64: 64 tet cubes (the default bench line); hybrid: 512 affine "
                         "boxes fired into a 983k-tet bed; cloth_abd: 256x256 cloth over 16 "
                         "affine boxes")
    ap.add_argument("--settle", type=int, default=None,
                    help="untimed time steps before warmup (default 40 / 4 / 60 by workload)")
    ap.add_argument("--n-side", type=int, default=4)
    ap.add_argument("--cells", type=int, default=8)
    ap.add_argument("--mas-fp64", action="store_true", help="all-FP64 MAS instead of FP32 apply")
    ap.add_argument("--cpu-iters", type=int, default=1, help="oracle Newton iterations timed")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="all", choices=["all", "ref"],
                    help="time all solver modes or only the reference FP64 one")
    ap.add_argument("--train-iters", type=int, default=12,
                    help="Newton solves whose coarse systems train the network")
    ap.add_argument("--train-epochs", type=int, default=400)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is design-only", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import sgipc
    from sgipc import scene as S

    log("device:", sgipc.device_info())
    if args.workload == "hybrid":
        sc = S.hybrid_scene()
        settle_default = 4
        workload = (f"512 affine boxes (6^3 cells, 1,296 tets each, kappa 1e11) at 10 m/s into "
                    f"a 64x64x40-cell bed (983,040 tets, E 5e4 Pa, bottom layer fixed); "
                    f"{sc.n_tets} tets, {sc.n_verts} verts; dt 0.005, dhat 5e-4, mu 0.3")
    elif args.workload == "cloth_abd":
        sc = S.cloth_scene(n=256, size=1.0, height=0.3, n_boxes=16, box_size=0.1, box_cells=3,
                           E=1e6, bend_stiffness=1e-4, thickness=1e-3, abd_boxes=True,
                           abd_kappa=1e11)
        settle_default = 60
        workload = (f"cloth sheet 256x256 verts ({sc.tris.shape[0]} triangles, 1 m), membrane "
                    f"E 1e6 Pa, bending 1e-4, thickness 1e-3, dropped 0.3 m over 16 affine boxes "
                    f"(kappa 1e11); dt 0.01, dhat 1e-3")
    else:
        sc = S.cubes_scene(n_side=args.n_side, cells=args.cells, size=0.2, spacing=0.6, fall=0.5,
                           seed=0)
        settle_default = 40
        workload = None
    if args.settle is None:
        args.settle = settle_default
    sc.mas_precision = sgipc.SGIPC_MAS_FP64 if args.mas_fp64 else sgipc.SGIPC_MAS_FP32
    sc.pcg_rel_tol = 1e-4
    nparts = int(sc.part_id.max()) + 1
    log(f"scene: {sc.n_verts} verts, {sc.n_tets} tets, {sc.faces.shape[0]} faces, "
        f"{sc.edges.shape[0]} edges, {nparts} banks")
    t0 = time.time()
    sim = sgipc.Simulation(sc)
    log(f"create: {time.time() - t0:.2f} s")

    # ---- settle (untimed): integrate until the lowest layer is in contact
    t0 = time.time()
    for k in range(args.settle):
        st = sim.time_step()
        if k % 10 == 9 or k == args.settle - 1 or args.settle <= 10 or args.workload != "cubes64":
            x = sim.positions()
            log(f"settle step {k + 1}/{args.settle}: newton={st['newton_iterations']} "
                f"pcg={st['pcg_iterations']} ymin={x[:, 1].min():.4f} "
                f"({time.time() - t0:.1f} s)")
    state = (sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)

    # ---- training data for the learned coarse correction (untimed): solves
    # of the settled scene with the affine warm start, tight tolerance so the
    # targets carry the coupling between aggregates; full-batch Adam on GPU
    train = None
    if args.modes == "all" and args.train_iters > 0:
        t0 = time.time()
        sim.set_pcg(rel_tol=1e-8, use_rz=True, solver_precision=sgipc.SGIPC_SOLVE_FP64)
        sim.learned_mode(1)
        F, T = [], []
        for _ in range(args.train_iters):
            sim.advance(1)
            f, t = sim.learned_samples()
            F.append(f)
            T.append(t)
        F, T = np.concatenate(F), np.concatenate(T)
        t1 = time.time()
        loss = sim.mlp_train(F, T, epochs=args.train_epochs, lr=1e-3)
        t2 = time.time()
        sim.learned_mode(0)
        train = {"samples": int(F.shape[0]), "collect_s": round(t1 - t0, 2),
                 "train_s": round(t2 - t1, 2), "epochs": args.train_epochs,
                 "loss_untrained": float((T.astype(np.float64) ** 2).mean()),
                 "loss_trained": loss}
        log(f"learned correction: {train}")

    # ---- timed Newton iterations, each solver mode from the same settled state.
    # fp64_ref (the value): the reference's plain FP64 solve, stop test on the
    # preconditioned norm r.z every 10 iterations.  fp64_true / mixed_true: stop
    # on the FP64 true residual ||b - A x|| / ||b|| <= 1e-4, FP64 PCG versus FP32
    # inner PCG with FP64 refinement.
    # fp64_affine / fp64_learned / mixed_learned: the same solves continued
    # from the affine coarse guess, or from the guess the trained network
    # corrects (MAS picks the better guess; zero-guess fallback), with the
    # zero-guess stop threshold.
    ref = dict(rel_tol=1e-4, use_rz=True, solver_precision=sgipc.SGIPC_SOLVE_FP64)
    true64 = dict(rel_tol=1e-4, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_FP64)
    true32 = dict(rel_tol=1e-4, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_MIXED)
    modes = [
        ("fp64_ref", ref, 0),
        ("fp64_true", true64, 0),
        ("mixed_true", true32, 0),
        ("fp64_affine", ref, 1),
        ("fp64_learned", ref, 2),
        ("mixed_learned", true32, 2),
    ]
    results = {}
    for name, cfg, lmode in modes:
        if args.modes != "all" and name != "fp64_ref":
            continue
        if lmode == 2 and train is None:
            continue
        sim.set_state(*state)
        sim.set_pcg(**cfg)
        sim.learned_mode(lmode)
        if args.warmup:
            sim.advance(args.warmup)
        # HIP-event timing of every 10th launch of the profiled kernels during
        # the timed region (one event pair adds a few microseconds of stream time)
        sim.profile(10 if name == "fp64_ref" else 0)
        t0 = time.perf_counter()
        adv = sim.advance(args.steps)
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        results[name] = {"ms_per_newton": round(ms, 4),
                         "pcg_iterations_per_newton": round(adv["pcg_iterations"] / args.steps, 2),
                         "max_rel_residual": adv["max_pcg_rel_residual"],
                         "fp64_fallbacks": adv["solver_fallbacks"]}
        if lmode:
            results[name]["warm_starts"] = adv["warm_starts"]

        log(f"timed [{name}]: {args.steps} Newton iterations in {(t1 - t0):.3f} s -> "
            f"{ms:.3f} ms/iter; pcg/iter={adv['pcg_iterations'] / args.steps:.1f} "
            f"max_rel_res={adv['max_pcg_rel_residual']:.2e} "
            f"steps_begun={adv['time_steps_begun']} fallbacks={adv['solver_fallbacks']}")
        if name == "fp64_ref":
            kern = {}
            for kname in ("spmv", "mas_apply_l0", "fem", "mas_invert"):
                n, tot = sim.profile_query(kname)
                if n:
                    kern[kname] = (n, tot)
                    log(f"kernel {kname}: {n} launches, {tot:.3f} ms total, "
                        f"{tot / n * 1e3:.2f} us avg")
            ref_adv, ref_ms = adv, ms
            sim.profile(0)
        # phase split (untimed, after the timed region): device time of
        # assembly + stop test, MAS build + solve, CCD bounds, line search,
        # averaged over a few further Newton iterations that solve
        ph = {"assembly": 0.0, "solve": 0.0, "ccd": 0.0, "line_search": 0.0}
        nph = 0
        for _ in range(6):
            it = sim.newton_iteration()
            if it["converged"]:
                sim.advance(1)  # begins the next time step
                continue
            for key in ph:
                ph[key] += it["ms_" + key]
            nph += 1
            if nph == 3:
                break
        if nph:
            results[name]["phases_ms"] = {k: round(v / nph, 3) for k, v in ph.items()}
            log(f"phases [{name}]: {results[name]['phases_ms']}")
    sim.learned_mode(0)
    adv, ms = ref_adv, ref_ms
    sz = sim.sizes()
    log(f"sizes={sz}")

    # ---- roofline of the dominant profiled kernel (HIP events on the library stream)
    hbm_kernels = {k: v for k, v in kern.items() if k in ("spmv", "mas_apply_l0", "fem")}
    dom = max(hbm_kernels, key=lambda k: hbm_kernels[k][1]) if hbm_kernels else None
    roof = None
    if dom:
        n, tot = kern[dom]
        avg_s = tot / n * 1e-3
        byt = algorithmic_bytes(dom, sz, nparts)
        ach = byt / avg_s / 1e9
        default_workload = (args.workload, args.n_side, args.cells, args.settle) == (
            "cubes64", 4, 8, 40)
        traffic = measured_traffic(dom) if default_workload else None
        roof = {"bound": "hbm", "kernel": dom, "achieved": round(ach, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": "profiles/spmv_64cubes_hbm_traffic.md" if traffic else None,
                "algorithmic_bytes_per_launch": int(byt), "avg_launch_us": round(avg_s * 1e6, 3),
                "launches": n}

    # ---- CPU baseline: the oracle (single thread) from the same settled state
    cpu = None
    if not args.no_cpu and args.cpu_iters > 0 and args.workload == "cubes64":
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import pyoracle as O
            w = O.World(sc)
            w.set_state(*state)
            c0 = time.perf_counter()
            w.advance(args.cpu_iters)
            c1 = time.perf_counter()
            cms = (c1 - c0) * 1e3 / args.cpu_iters
            cpu = {"value": round(cms, 3), "unit": "ms/Newton step", "cores": 1, "kind": "port",
                   "sample": f"{args.cpu_iters} Newton iterations of the same 64-cube scene "
                             f"from the settled state (oracle, FP64, single thread)"}
            log(f"cpu baseline: {cms:.1f} ms/iter")
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
        "dtype": "f64",
        "data": "synthetic",
        "pcg_iterations_per_newton": round(adv["pcg_iterations"] / args.steps, 2),
        "max_pcg_rel_residual": adv["max_pcg_rel_residual"],
        "modes": results,
        "learned_training": train,
        "config": {
            "workload": (workload or
                         f"{args.n_side ** 3} tet cubes of {args.cells}^3 hex cells "
                         f"({sc.n_tets} tets, {sc.n_verts} verts), E log-uniform [1e5,1e9] Pa, "
                         f"nu 0.4, 0.5 m drop onto ground, dt 0.01, dhat 1e-3")
                        + f", {args.settle} settle steps",
            "mode": "FP64 PCG/SpMV, " + ("FP64" if args.mas_fp64 else "FP32") + " MAS apply",
            "pcg_rel_tol": 1e-4,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
