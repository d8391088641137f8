"""Scene-file layer of the reference CLI (host code): Gmsh reader pinned by
MshIO's own fixtures (tests/golden/msh: the same 2-quad mesh as 2.2/4.1,
ASCII/binary, MshHelper/MshIO/data), the tet-mesh view and surface of the
example meshes (tests/golden/scenes, from examples/) against closed forms of
a 75x5x5-cell box, .part files and the MAS bank maps against a sequential
restatement of setMAS_partition (main.cu:47-75), and the JSON schema
(docs/cli.md) on the reference's example scenes."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BAR = os.path.join(GOLD, "scenes", "bar_width_15_1_1_num_75_5_5_sorted.16_sorted.16")


@pytest.mark.parametrize("name", ["test_2.2_ascii", "test_2.2_bin", "test_4.1_ascii",
                                  "test_4.1_bin"])
def test_msh_fixtures_parse_to_the_same_mesh(name):
    from sgipc import msh
    m = msh.read_msh(os.path.join(GOLD, "msh", name + ".msh"))
    assert m.num_nodes == 6 and m.num_elements == 2
    xyz = np.concatenate([b.xyz for b in m.node_blocks if b.tags.size])
    tags = np.concatenate([b.tags for b in m.node_blocks if b.tags.size])
    np.testing.assert_array_equal(tags, np.arange(1, 7))
    np.testing.assert_array_equal(xyz, [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [2, 0, 0],
                                        [2, 1, 0]])
    (blk,) = m.element_blocks
    assert blk.element_type == 3  # 4-node quadrangles
    np.testing.assert_array_equal(blk.tags, [1, 2])
    np.testing.assert_array_equal(blk.nodes, [[1, 2, 3, 4], [2, 5, 6, 3]])
    with pytest.raises(ValueError, match="No tetrahedra"):
        msh.load_tet_mesh(os.path.join(GOLD, "msh", name + ".msh"))


def test_ground_plane_tet_mesh():
    from sgipc import msh
    V, T = msh.load_tet_mesh(os.path.join(GOLD, "scenes", "ground_plane_sorted.16.msh"))
    np.testing.assert_array_equal(V, [[-1, 0, -1], [1, 0, -1], [1, 0, 1], [-1, 0, 1],
                                      [-1, .1, -1], [1, .1, -1], [1, .1, 1], [-1, .1, 1]])
    np.testing.assert_array_equal(T, [[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6],
                                      [0, 4, 5, 6], [0, 5, 1, 6]])


def test_bar_mesh_surface_closed_forms():
    """75 x 5 x 5 hex cells of 6 tets: 11,250 tets; boundary quads 2 (75*5 +
    75*5 + 5*5) = 1,550 -> 3,100 triangles; interior nodes 74*4*4 -> 1,552
    surface vertices; Euler V - E + F = 2 -> 4,650 edges; every face normal
    points out of the (convex) bar"""
    from sgipc import msh
    from sgipc.scene_json import tet_surface
    V, T = msh.load_tet_mesh(BAR + ".msh")
    assert V.shape == (76 * 6 * 6, 3) and T.shape == (11250, 4)
    vol = np.einsum("ij,ij->i", np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]]),
                    V[T[:, 3]] - V[T[:, 0]]) / 6
    assert np.abs(vol).sum() == pytest.approx(15.0, rel=1e-12)
    faces, edges, sv = tet_surface(V, T)
    assert (faces.shape[0], edges.shape[0], sv.size) == (3100, 4650, 1552)
    p = V[faces]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    c = p.mean(1) - V.mean(0)
    assert np.all(np.einsum("ij,ij->i", n, c) > 0)


def _set_mas_partition_loop(part):
    """setMAS_partition (main.cu:47-75) as its loops read"""
    nparts = int(part[-1]) + 1
    mp = [-1] * (nparts * 16)
    remap = [0] * len(part)
    index = 0
    for i in range(len(part)):
        mp[16 * part[i] + index] = i
        index += 1
        if i <= len(part) - 2 and part[i + 1] != part[i]:
            index = 0
    index = 0
    for i in range(len(mp)):
        if mp[i] == index:
            remap[index] = i
            index += 1
    return np.array(mp), np.array(remap)


def test_part_file_and_mas_maps():
    from sgipc.scene_json import load_part_file, mas_partition_maps
    ids, nxt = load_part_file(BAR + ".part", 2736, 5)
    assert ids.size == 2736 and ids[0] >= 5 and nxt == ids[-1] + 1
    assert np.all(np.diff(ids) >= 0) and np.bincount(ids - 5).max() <= 16
    mp, remap = mas_partition_maps(ids - 5)
    mpl, remapl = _set_mas_partition_loop(list(ids - 5))
    np.testing.assert_array_equal(mp, mpl)
    np.testing.assert_array_equal(remap, remapl)
    np.testing.assert_array_equal(mp[remap], np.arange(2736))
    with pytest.raises(ValueError, match="lost the data"):
        load_part_file(BAR + ".part", 2735, 0)


def test_free_fall_scene():
    from sgipc.scene_json import load_scene_json
    run = load_scene_json(os.path.join(GOLD, "scenes", "free_fall.json"))
    sc = run.scene
    assert (run.frames, run.preconditioner_type, run.write_obj_frames) == (100, 0, True)
    assert run.object_ranges == [(0, 8), (8, 8 + 2736)]
    assert sc.n_verts == 2744 and sc.n_tets == 6 + 11250
    assert np.all(sc.btype[:8] == 1) and np.all(sc.btype[8:] == 0)
    bar = sc.x[8:]
    # scale 0.1 about the box centre (7.5, .5, .5), then translate (-0.75, 1, -0.05)
    np.testing.assert_allclose(bar.min(0), [-1.5, 0.95, -0.1], atol=1e-15)
    np.testing.assert_allclose(bar.max(0), [0.0, 1.05, 0.0], atol=1e-15)
    # dhat from the largest-volume box: the ground plane's 2 x 0.1 x 2
    assert sc.bbox_diag == pytest.approx(np.sqrt(8.01), rel=1e-15)
    assert sc.dhat == pytest.approx(1e-3 * np.sqrt(8.01), rel=1e-15)
    assert sc.scene_diag == pytest.approx(np.sqrt(1.5 ** 2 + 0.02), rel=1e-12)
    assert (sc.dt, sc.friction_mu, sc.use_ground, sc.precond) == (0.02, 0.2, False, 0)
    assert sc.pcg_rel_tol == 0.01 and sc.pcg_use_preconditioned_norm == 1
    assert sc.gravity == (0.0, -9.8, 0.0) and sc.soft_index is None


def _bar_wall_json(tmp_path, **extra):
    d = json.load(open(os.path.join(GOLD, "scenes", "bar_on_the_wall.json")))
    d["objects"][0]["mesh_msh"] = BAR + ".msh"
    d["objects"][0]["part_file"] = BAR + ".part"
    d.update(extra)
    p = tmp_path / "bar_wall.json"
    p.write_text(json.dumps(d))
    return p


def test_pins_parts_and_animation(tmp_path):
    from sgipc.scene_json import load_scene_json
    # centred on its box the bar spans x in [-7.5, 7.5] with nodes every 0.2,
    # so the file's pin box x in [0, 0] selects no vertex
    run = load_scene_json(_bar_wall_json(tmp_path))
    assert int(run.scene.btype.sum()) == 0
    anim = [{"object": 0, "boxes": [{"min": [7.4, -1, -1], "max": [7.5, 1, 1]}],
             "rot_origin": [0, 0, 0], "rot_axis": [2, 0, 0], "rot_velocity": 1.5}]
    d = json.load(open(_bar_wall_json(tmp_path)))
    d["objects"][0]["pin_boxes"] = [{"min": [-7.5, -100, -100], "max": [-7.5, 100, 100]}]
    d["animation"] = anim
    (tmp_path / "b2.json").write_text(json.dumps(d))
    run = load_scene_json(tmp_path / "b2.json")
    sc = run.scene
    # the x = -7.5 end face (6 x 6 nodes) is pinned
    assert int(sc.btype.sum()) == 36 and np.all(sc.x[sc.btype == 1, 0] == -7.5)
    assert sc.precond == 1 and np.all(np.diff(sc.part_id) >= 0)
    # the x = 7.5 end face is animated
    assert sc.soft_index.size == 36 and np.all(sc.x[sc.soft_index, 0] == 7.5)
    np.testing.assert_array_equal(sc.soft_target, sc.x[sc.soft_index])
    np.testing.assert_array_equal(run.soft_rotations[0].axis, [1.0, 0.0, 0.0])
    t = run.animate_targets(sc.x)
    ang = sc.dt * 1.5
    np.testing.assert_allclose(t[:, 1], sc.x[sc.soft_index, 1] * np.cos(ang)
                               - sc.x[sc.soft_index, 2] * np.sin(ang), atol=1e-15)
    d["animation"] = anim + [dict(anim[0])]
    (tmp_path / "b3.json").write_text(json.dumps(d))
    with pytest.raises(ValueError, match="multiple animation constraints"):
        load_scene_json(tmp_path / "b3.json")


def test_missing_mesh_is_reported():
    from sgipc.scene_json import load_scene_json
    with pytest.raises(FileNotFoundError, match="octopus_mid_sorted.16.msh"):
        load_scene_json(os.path.join(GOLD, "scenes", "octopus_stack.json"))
