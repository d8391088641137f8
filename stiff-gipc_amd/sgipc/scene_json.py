"""Scene files of the reference CLI: the JSON schema (StiffGIPC/main.cu:
130-308, 649-880; docs/cli.md:77-442), Gmsh tet meshes, METIS ``.part``
files and the MAS bank maps built from them.

load_scene_json() turns one scene file into a Scene (the sgipc_scene_desc
the library takes) plus the run controls the CLI loop needs:

  settings   -> dt, PCG and Newton tolerances, Armijo constants, friction,
                gravity, motion stiffness / rate, dhat = IPC_ralative_dHat *
                the diagonal of the largest-volume object box
                (GIPC.cu:8536-8543 with main.cu's maxConer/minConer)
  objects[]  -> each mesh transformed about its box centre (rotate XYZ in
                degrees, scale, translate; load_mesh.cpp:483-596), its
                Young's modulus / density / Poisson ratio per tet, obstacles
                and pin boxes fixed (boundary type 1), initial velocity, and
                with preconditioner_type != 0 its .part file offset by the
                parts before it (tetrahedra_obj::load_parts,
                load_mesh.cpp:449-480)
  ground     -> plane normal / offset
  animation  -> soft targets: the vertices of an object inside its boxes, each
                in at most one constraint, rotated every frame about an axis
                (main.cu:805-851, 1116-1140)
  simulation -> frames, preconditioner type, OBJ output, frames whose linear
                systems are exported

Extensions beyond the reference schema (optional keys):
  objects[].body_type "ABD" makes the object an affine body (its tets give
  mass only); settings.abd_kappa sets the affine shape stiffness;
  simulation.mas_precision "fp64" keeps the MAS inverses and apply in FP64
  (the reference's MAS applies FP32 inverses: "fp32", the default).
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

import numpy as np

from .msh import load_tet_mesh
from .scene import Scene

BANKSIZE = 16


# --------------------------------------------------------------------------
# part files and the MAS bank maps
def load_part_file(path, n_verts, offset):
    """part ids of one object (whitespace-separated integers) shifted by
    offset; returns (ids, next offset = last id + 1), as
    tetrahedra_obj::load_parts does (load_mesh.cpp:449-480)"""
    with open(path) as f:
        ids = np.array([int(t) for t in f.read().split()], np.int64)
    if ids.size != n_verts:
        raise ValueError(f"part info file {path} lost the data: {ids.size} ids for "
                         f"{n_verts} vertices")
    ids = ids + offset
    return ids, int(ids[-1]) + 1


def mas_partition_maps(part_id):
    """(partId_map_real, real_map_partId) of setMAS_partition (main.cu:47-75):
    map[16 * part + k] = the k-th vertex of part `part` (-1 padding) and
    remap[v] = the position of vertex v in that padded layout.  The parts
    must be contiguous runs of at most 16 vertices."""
    p = np.asarray(part_id, np.int64)
    n = p.size
    nparts = int(p[-1]) + 1 if n else 0
    if n and (np.any(np.diff(p) < 0)):
        raise ValueError("part ids must be non-decreasing in vertex order")
    counts = np.bincount(p, minlength=nparts)
    if counts.size and counts.max() > BANKSIZE:
        raise ValueError(f"a part holds more than {BANKSIZE} vertices")
    mp = np.full(nparts * BANKSIZE, -1, np.int64)
    first = np.r_[0, np.cumsum(counts)[:-1]]
    k = np.arange(n) - first[p]
    mp[BANKSIZE * p + k] = np.arange(n)
    remap = np.empty(n, np.int64)
    remap[mp[mp >= 0]] = np.nonzero(mp >= 0)[0]
    return mp, remap


# --------------------------------------------------------------------------
# surfaces
def tet_surface(x, tets):
    """Boundary triangles oriented with the outward normal (the tet's fourth
    vertex behind the face, getSurface load_mesh.cpp:1003-1100), their
    unique edges and vertices; sorted for a deterministic order."""
    t = np.asarray(tets, np.int64)
    f = np.concatenate([t[:, [0, 1, 2]], t[:, [1, 2, 3]], t[:, [2, 3, 0]], t[:, [3, 0, 1]]])
    opp = np.concatenate([t[:, 3], t[:, 0], t[:, 1], t[:, 2]])
    key = np.sort(f, 1)
    _, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    keep = idx[cnt == 1]
    faces, o = f[keep], opp[keep]
    n = np.cross(x[faces[:, 1]] - x[faces[:, 0]], x[faces[:, 2]] - x[faces[:, 0]])
    flip = np.einsum("ij,ij->i", n, x[o] - x[faces[:, 0]]) > 0
    faces[flip] = faces[flip][:, [0, 2, 1]]
    faces = faces[np.lexsort((faces[:, 2], faces[:, 1], faces[:, 0]))]
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]])
    e = np.unique(np.sort(e, 1), axis=0)
    sv = np.unique(faces.ravel())
    return faces.astype(np.uint32), e.astype(np.uint32), sv.astype(np.uint32)


# --------------------------------------------------------------------------
# transforms
def euler_xyz(deg):
    """R = Rz Ry Rx of XYZ Euler angles in degrees (read_transform,
    main.cu:130-152; load_mesh.cpp AngleAxis product)"""
    ax, ay, az = np.deg2rad(np.asarray(deg, np.float64))
    cx, sx, cy, sy, cz, sz = np.cos(ax), np.sin(ax), np.cos(ay), np.sin(ay), np.cos(az), np.sin(az)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    Rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return Rz @ Ry @ Rx


def transform_vertices(V, rotation_deg, scale, translation):
    """rotate about the box centre, scale, translate (load_mesh.cpp:520-541)"""
    V = np.asarray(V, np.float64)
    c = (V.min(0) + V.max(0)) * 0.5
    R = euler_xyz(rotation_deg)
    return ((V - c) @ R.T) * float(scale) + np.asarray(translation, np.float64)


# --------------------------------------------------------------------------
@dataclass
class SoftRotation:
    origin: np.ndarray
    axis: np.ndarray          # unit
    velocity: float           # radians per second


@dataclass
class SceneRun:
    scene: Scene
    frames: int
    write_obj_frames: bool
    export_linear_system_frames: list
    object_ranges: list                         # (v_begin, v_end) per object
    soft_rotations: list = field(default_factory=list)  # one per soft target
    preconditioner_type: int = 0

    def animate_targets(self, x):
        """soft-target positions for the next frame from the current vertex
        positions x: each rotated about its constraint's axis by velocity * dt
        (main.cu:1116-1140)"""
        from .scene import rotate_axis_angle
        sc = self.scene
        out = np.empty((len(self.soft_rotations), 3))
        for i, (v, r) in enumerate(zip(sc.soft_index, self.soft_rotations)):
            out[i] = rotate_axis_angle(x[v][None, :], r.origin, r.axis, r.velocity * sc.dt)[0]
        return out


def _vec3(a, what):
    a = np.asarray(a, np.float64)
    if a.shape != (3,):
        raise ValueError(f"Expected {what} vec3")
    return a


def _resolve(path, json_dir):
    cands = [path, os.path.join(json_dir, path), os.path.join(json_dir, os.path.basename(path))]
    for c in cands:
        if os.path.exists(c):
            return c
    raise FileNotFoundError(f"scene file references {path}, found at none of {cands}")


def _in_box(x, bmin, bmax):
    return np.all((x >= bmin) & (x <= bmax), axis=1)


def load_scene_json(path):
    """Parse a reference scene file (every key of the reference schema is
    required, as json.at(...) requires it) into a SceneRun."""
    with open(path) as f:
        j = json.load(f)
    json_dir = os.path.dirname(os.path.abspath(path))
    s, sim, ground = j["settings"], j["simulation"], j["ground"]
    pre = int(sim["preconditioner_type"])
    xs, ts, Es, nus, rhos, bts, vels, parts, abd_flags = [], [], [], [], [], [], [], [], []
    ranges = []
    part_offset = 0
    nv = 0
    best_vol, best_box = -1.0, None
    non_obs = [np.full(3, np.inf), np.full(3, -np.inf)]
    for obj in j["objects"]:
        mesh = _resolve(obj["mesh_msh"], json_dir)
        part_file = obj["part_file"]
        is_obstacle = bool(obj["is_obstacle"])
        E, rho, nu = float(obj["young_modulus"]), float(obj["density"]), float(obj["poisson_ratio"])
        tr = obj["transform"]
        rot = _vec3(tr["rotation"], "transform.rotation")
        trans = _vec3(tr["translation"], "transform.translation")
        v0 = _vec3(obj["initial_velocity"], "initial_velocity")
        body_type = obj.get("body_type", "FEM")
        V, T = load_tet_mesh(mesh)
        X = transform_vertices(V, rot, float(tr["scale"]), trans)
        lo, hi = X.min(0), X.max(0)
        vol = float(np.prod(hi - lo))
        if vol > best_vol:  # the largest-volume object box (load_mesh.cpp:583-593)
            best_vol, best_box = vol, (lo, hi)
        if not is_obstacle:
            non_obs = [np.minimum(non_obs[0], lo), np.maximum(non_obs[1], hi)]
        n = X.shape[0]
        bt = np.full(n, 1 if is_obstacle else 0, np.int32)
        for box in obj["pin_boxes"]:
            bt[_in_box(X, _vec3(box["min"], "pin_boxes.min"), _vec3(box["max"], "pin_boxes.max"))] = 1
        if pre != 0:
            ids, part_offset = load_part_file(_resolve(part_file, json_dir), n, part_offset)
            parts.append(ids)
        xs.append(X)
        ts.append(T + nv)
        Es.append(np.full(T.shape[0], E))
        nus.append(np.full(T.shape[0], nu))
        rhos.append(np.full(T.shape[0], rho))
        bts.append(bt)
        vels.append(np.tile(v0, (n, 1)))
        abd_flags.append(body_type == "ABD")
        ranges.append((nv, nv + n))
        nv += n
    x = np.concatenate(xs)
    tets = np.concatenate(ts)
    btype = np.concatenate(bts)
    # affine bodies come after every FEM vertex (sgipc_scene_desc convention)
    order = [i for i, a in enumerate(abd_flags) if not a] + [i for i, a in enumerate(abd_flags) if a]
    if order != list(range(len(order))):
        raise ValueError("affine-body objects must follow every FEM object in the scene file")
    body_id = np.full(nv, -1, np.int32)
    abd_body = None
    if any(abd_flags):
        abd_body = np.full(nv, -1, np.int32)
        b = 0
        for (a, (v0_, v1_)) in zip(abd_flags, ranges):
            if a:
                abd_body[v0_:v1_] = b
                body_id[v0_:v1_] = b
                b += 1
    if pre != 0:
        part_id = np.concatenate(parts).astype(np.int32)
        mas_partition_maps(part_id)  # validates contiguous parts of <= 16
    else:
        part_id = (np.arange(nv) // BANKSIZE).astype(np.int32)
    faces, edges, surf = tet_surface(x, tets)
    lo, hi = best_box
    bbox_diag = float(np.linalg.norm(hi - lo))
    scene_diag = float(np.linalg.norm(non_obs[1] - non_obs[0])) if np.isfinite(non_obs[0]).all() \
        else 0.0
    sc = Scene(
        x=x, tets=tets.astype(np.uint32), E=np.concatenate(Es), nu=np.concatenate(nus),
        rho=np.concatenate(rhos), part_id=part_id, body_id=body_id, btype=btype,
        faces=faces, edges=edges, surf_verts=surf, v=np.concatenate(vels),
        abd_body=abd_body, abd_kappa=float(s.get("abd_kappa", 1e8)),
        dt=float(s["ipc_time_step"]), dhat=float(s["IPC_ralative_dHat"]) * bbox_diag,
        gravity=tuple(_vec3(s["gravity"], "gravity")), use_ground=bool(ground["enabled"]),
        ground_normal=tuple(_vec3(ground["normal"], "ground.normal")),
        ground_offset=float(ground["offset"]), friction_mu=float(s["friction_rate"]),
        ground_friction_mu=float(s["gd_friction_rate"]),
        abs_xdelta_tol=float(s["abs_xdelta_tol"]), rel_xdelta_tol=float(s["rel_xdelta_tol"]),
        armijo_c1=float(s["armijo_c1"]), armijo_beta=float(s["armijo_beta"]),
        armijo_alpha_min=float(s["armijo_alpha_min"]),
        pcg_rel_tol=float(s["pcg_rel_threshold"]), pcg_abs_tol=float(s["pcg_abs_threshold"]),
        pcg_max_iter=int(s["pcg_max_iter"]),
        pcg_use_preconditioned_norm=int(bool(s["pcg_use_preconditioned_norm"])),
        precond=1 if pre != 0 else 0, soft_stiffness=float(s["motion_stiffness"]),
        mas_precision=1 if sim.get("mas_precision", "fp32") == "fp64" else 0,
        motion_rate=float(s["motion_rate"]), bbox_diag=bbox_diag, scene_diag=scene_diag,
        name=os.path.splitext(os.path.basename(path))[0])
    # the settings the reference reads but the CLI does not use (cloth keys,
    # legacy density / Poisson ratio, buffer scale) must still be present
    for k in ("volume_mesh_density", "poisson_rate", "triangle_mesh_thickness",
              "triangle_mesh_youngs_modulus", "triangle_bend_youngs_modulus",
              "triangle_mesh_density", "strain_rate", "collision_detection_buff_scale"):
        s[k]
    # animated soft constraints (read_animation main.cu:237-262; box selection
    # main.cu:805-851)
    soft_idx, rots = [], []
    taken = np.zeros(nv, bool)
    for c in j["animation"]:
        o = int(c["object"])
        if o < 0 or o >= len(ranges):
            raise ValueError("animation.constraints[].object out of range")
        a, b = ranges[o]
        axis = _vec3(c["rot_axis"], "rot_axis")
        axis = axis / np.linalg.norm(axis)
        rot = SoftRotation(_vec3(c["rot_origin"], "rot_origin"), axis, float(c["rot_velocity"]))
        for box in c["boxes"]:
            sel = a + np.nonzero(_in_box(x[a:b], _vec3(box["min"], "min"),
                                         _vec3(box["max"], "max")))[0]
            for v in sel:
                if taken[v]:
                    raise ValueError("A vertex is selected by multiple animation constraints")
                taken[v] = True
                soft_idx.append(int(v))
                rots.append(rot)
    if soft_idx:
        sc.soft_index = np.asarray(soft_idx, np.uint32)
        sc.soft_target = x[sc.soft_index].copy()
    sc.meta = {"json": os.path.abspath(path)}
    return SceneRun(scene=sc, frames=int(sim["frames"]),
                    write_obj_frames=bool(sim["write_obj_frames"]),
                    export_linear_system_frames=[int(f) for f in
                                                 sim["export_linear_system_frames"]],
                    object_ranges=ranges, soft_rotations=rots, preconditioner_type=pre)
