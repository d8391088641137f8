"""The reference CLI contract end to end on the GPU (main.cu:649-1163): a
scene file from the reference's examples (a 75x5x5-cell bar, MAS banks from
its .part file, one end pinned, the other end animated by rotating soft
targets) runs frame by frame; stats.json, the OBJ frames and the exported
linear system are written; positions match the oracle stepping the same
loaded scene to 1e-8 per frame, and the exported system solves in the
solve-only mode as the oracle's PCG on the same file does."""
import json
import os

import numpy as np
import pytest
import scipy.io

from helpers import rel

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BAR = os.path.join(GOLD, "scenes", "bar_width_15_1_1_num_75_5_5_sorted.16_sorted.16")


def _scene_file(tmp_path):
    d = json.load(open(os.path.join(GOLD, "scenes", "bar_on_the_wall.json")))
    o = d["objects"][0]
    o["mesh_msh"], o["part_file"] = BAR + ".msh", BAR + ".part"
    o["pin_boxes"] = [{"min": [-7.5, -100, -100], "max": [-7.5, 100, 100]}]
    d["animation"] = [{"object": 0, "boxes": [{"min": [7.4, -1, -1], "max": [7.5, 1, 1]}],
                       "rot_origin": [7.5, 0, 0], "rot_axis": [1, 0, 0], "rot_velocity": 0.5}]
    d["settings"]["motion_stiffness"] = 1e6
    d["simulation"]["export_linear_system_frames"] = [1]
    # FP64 MAS (an extension key) and a tight PCG threshold: with the file's
    # pcg_rel_threshold 1e-2, the PCG stops on an iterate that rounding steers
    # (FP64 atomics order on the GPU, sequential sums in the oracle), and the
    # two trajectories part at 1e-7 within 3 frames
    d["simulation"]["mas_precision"] = "fp64"
    d["settings"]["pcg_rel_threshold"] = 1e-9
    p = tmp_path / "scene.json"
    p.write_text(json.dumps(d))
    return p


def test_cli_run_matches_oracle(tmp_path, oracle):
    from sgipc import cli
    from sgipc.scene_json import load_scene_json
    path = _scene_file(tmp_path)
    out = tmp_path / "out"
    sim, stats = cli.run(str(path), str(out), frames=3)
    run = load_scene_json(str(path))
    w = oracle.World(run.scene)
    for frame in range(3):
        w.set_soft_targets(run.animate_targets(w.positions()))
        so = w.time_step()
        print(f"[scene] frame={frame} oracle newton={so['newton_iterations']} "
              f"gpu newton={len(stats['frames'][frame]['newton'])}", flush=True)
    err = rel(sim.positions(), w.positions())
    print(f"[scene] positions after 3 frames rel={err:.2e}", flush=True)
    assert err < 1e-8
    # stats.json: one entry per frame, every Newton iteration, the last one the
    # stop test (no PCG); a GlobalTimer tree in seconds
    st = json.load(open(out / "stats.json"))
    assert len(st["frames"]) == 3
    for f in st["frames"]:
        assert len(f["newton"]) >= 2 and "pcg" not in f["newton"][-1]
        assert all(e["pcg"]["iterations"] > 0 for e in f["newton"][:-1])
        assert f["timer"]["name"] == "GlobalTimer" and f["timer"]["parent"] == ""
        assert f["timer"]["children"][0]["name"] == "IPC_Solver"
    # OBJ frames 0..3 with every vertex and the oriented surface
    for k in range(4):
        lines = open(out / f"frame_{k:05d}.obj").read().split("\n")
        assert sum(l.startswith("v ") for l in lines) == 2736
        assert sum(l.startswith("f ") for l in lines) == 3100
    # frame 1's systems, one per solving Newton iteration
    A_path = out / "linear_system" / "frame_00001_ns_0_A.mtx"
    b_path = out / "linear_system" / "frame_00001_ns_0_b.mtx"
    A = scipy.io.mmread(str(A_path)).tocsr()
    assert A.shape == (3 * 2736, 3 * 2736) and abs(A - A.T).max() == 0.0
    n_ns = len([f for f in os.listdir(out / "linear_system") if f.startswith("frame_00001_ns_")
                and f.endswith("_A.mtx")])
    assert n_ns == len(st["frames"][1]["newton"]) - 1
    # solve-only mode on it, against the oracle's PCG (MAS banks of the .part file)
    x, s1 = cli.solve_only(str(path), str(A_path), str(b_path))
    import sgipc
    nb, r, c, v = sgipc.read_mm_upper_bcoo(str(A_path))
    Ao = oracle.bsr_from_upper_triplets(nb, r, c, v.reshape(-1, 9))
    b = sgipc.read_mm_vector(str(b_path))
    m = oracle.Mas(Ao, run.scene.part_id, fp32=False)
    xo, ito, _, _ = oracle.pcg(Ao, b, m, 1, rel_tol=run.scene.pcg_rel_tol,
                               use_rz=run.scene.pcg_use_preconditioned_norm)
    print(f"[scene] solve-only iterations gpu={s1.iterations} oracle={ito}", flush=True)
    assert s1.iterations == ito and rel(x, xo) < 1e-8
