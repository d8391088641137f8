"""Headless driver with the reference CLI's contract (StiffGIPC/main.cu:
649-1163, docs/cli.md):

  python -m sgipc.cli --json scene.json -o out_dir
      runs simulation.frames time steps of the scene; writes
      out_dir/frame_00000.obj (initial state) and frame_<k>.obj after frame
      k - 1 when write_obj_frames, out_dir/stats.json after every frame
      (frames[i].newton[k].pcg.iterations and a timer tree in seconds), and
      the linear systems of the frames listed in export_linear_system_frames
      under out_dir/linear_system/.
  python -m sgipc.cli --json scene.json --A A.mtx --b b.mtx -o out_dir
      solve-only mode: the scene configures the solver (preconditioner, PCG
      settings); A and b are loaded from MatrixMarket files and solved once.

The --sort mode (offline METIS re-ordering) is outside this build.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def write_obj(path, x, faces):
    """vertices and surface triangles, 1-based (write_obj, main.cu:358-374)"""
    with open(path, "w") as f:
        for v in x:
            f.write(f"v {v[0]:.17g} {v[1]:.17g} {v[2]:.17g}\n")
        for t in faces:
            f.write(f"f {t[0] + 1} {t[1] + 1} {t[2] + 1}\n")


def _timer(name, parent, seconds, count, children=()):
    ch = sorted(children, key=lambda c: -c["duration"])
    return {"name": name, "duration": seconds, "count": count, "parent": parent, "children": ch}


def frame_stats(log, wall_s):
    """stats.json entry of one frame: the Newton list and a merged timer tree
    (GlobalTimer > IPC_Solver > newton_iteration > phases) built from the
    per-iteration device times"""
    newton = []
    for it in log:
        e = {}
        if not it["converged"]:
            e["pcg"] = {"iterations": int(it["pcg_iterations"])}
        newton.append(e)
    p_it = "/GlobalTimer/IPC_Solver/newton_iteration"
    phases = [("computeGradientAndHessian", "ms_assembly"), ("solve_linear_system", "ms_solve"),
              ("step_size_bounds", "ms_ccd"), ("lineSearch", "ms_line_search")]
    kids = [_timer(n, p_it, sum(it[k] for it in log) * 1e-3,
                   sum(1 for it in log if it[k] > 0 or k == "ms_assembly"))
            for n, k in phases]
    it_t = _timer("newton_iteration", "/GlobalTimer/IPC_Solver",
                  sum(it["ms_total"] for it in log) * 1e-3, len(log), kids)
    solver = _timer("IPC_Solver", "/GlobalTimer", wall_s, 1, [it_t])
    return {"newton": newton, "timer": _timer("GlobalTimer", "", wall_s, 1, [solver])}


def run(json_path, out_dir, frames=None, log=print):
    import sgipc
    from sgipc.scene_json import load_scene_json
    run_ = load_scene_json(json_path)
    sc = run_.scene
    nframes = run_.frames if frames is None else frames
    os.makedirs(out_dir, exist_ok=True)
    sim = sgipc.Simulation(sc)
    if run_.write_obj_frames:
        write_obj(os.path.join(out_dir, "frame_00000.obj"), sim.positions(), sc.faces)
    stats = {"frames": []}
    for frame in range(nframes):
        if frame in run_.export_linear_system_frames:
            sim.export_next_step(out_dir, frame)
        if sc.soft_index is not None and len(run_.soft_rotations):
            sim.set_soft_targets(run_.animate_targets(sim.positions()))
        t0 = time.perf_counter()
        st = sim.time_step()
        wall = time.perf_counter() - t0
        stats["frames"].append(frame_stats(sim.step_log(), wall))
        with open(os.path.join(out_dir, "stats.json"), "w") as f:
            json.dump(stats, f, indent=2)
        if run_.write_obj_frames:
            write_obj(os.path.join(out_dir, f"frame_{frame + 1:05d}.obj"), sim.positions(),
                      sc.faces)
        log(f"frame {frame}: newton={st['newton_iterations']} pcg={st['pcg_iterations']} "
            f"capped={st['capped']} {wall * 1e3:.1f} ms")
    return sim, stats


def solve_only(json_path, A_path, b_path, log=print):
    """main.cu:1047-1087: the scene's solver settings on a loaded system"""
    import sgipc
    from sgipc.scene_json import load_scene_json
    sc = load_scene_json(json_path).scene
    b = sgipc.read_mm_vector(b_path)
    nb, _, _, _ = sgipc.read_mm_upper_bcoo(A_path)
    if 3 * nb != b.size:
        raise ValueError(f"Loaded b has size={b.size} but expected M={3 * nb}")
    part = sc.part_id if (sc.precond == 1 and sc.part_id.size == nb) else None
    ls = sgipc.LinearSystem.from_matrix_market(A_path, part, sc.precond, sc.mas_precision)
    x, st = ls.pcg(b, rel_tol=sc.pcg_rel_tol, abs_tol=sc.pcg_abs_tol, max_iter=sc.pcg_max_iter,
                   use_rz=bool(sc.pcg_use_preconditioned_norm))
    log(f"Loaded linear system solved. Iterations: {st.iterations}")
    return x, st


def main(argv=None):
    ap = argparse.ArgumentParser(prog="sgipc")
    ap.add_argument("-j", "--json", required=True)
    ap.add_argument("--A", help="solve a provided MatrixMarket A (coordinate, real)")
    ap.add_argument("--b", help="... with the MatrixMarket b (array, real, M x 1)")
    ap.add_argument("-o", "--output", required=True)
    ap.add_argument("--frames", type=int, default=None, help="override simulation.frames")
    a = ap.parse_args(argv)
    if a.A or a.b:
        if not (a.A and a.b):
            print("When solving a provided linear system, both --A and --b are required.",
                  file=sys.stderr)
            return 1
        solve_only(a.json, a.A, a.b)
        return 0
    run(a.json, a.output, a.frames)
    return 0


if __name__ == "__main__":
    sys.exit(main())
