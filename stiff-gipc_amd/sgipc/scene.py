"""Synthetic scene construction and the ctypes mirror of ``sgipc_scene_desc``.

The reference loads sorted Gmsh meshes and METIS ``.part`` files
(StiffGIPC/load_mesh.cpp:449-596, MeshProcess/metis_partition) and gathers
settings from the scene JSON (StiffGIPC/main.cu:263-356).  This module builds
the same inputs synthetically: structured hex grids split into 6 tets per
cell, vertices ordered so that each 2x2x4 vertex block is one 16-vertex MAS
bank (part ids non-decreasing, as the reference's sorted meshes guarantee),
plus the collision surface (faces, edges, surface vertices).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field, replace

import numpy as np


class SceneDesc(C.Structure):
    """ctypes layout of include/sgipc_scene.h::sgipc_scene_desc."""

    _fields_ = [
        ("n_verts", C.c_int32),
        ("x", C.POINTER(C.c_double)),
        ("x_rest", C.POINTER(C.c_double)),
        ("v", C.POINTER(C.c_double)),
        ("btype", C.POINTER(C.c_int32)),
        ("body_id", C.POINTER(C.c_int32)),
        ("part_id", C.POINTER(C.c_int32)),
        ("n_tets", C.c_int32),
        ("tets", C.POINTER(C.c_uint32)),
        ("tet_E", C.POINTER(C.c_double)),
        ("tet_nu", C.POINTER(C.c_double)),
        ("tet_rho", C.POINTER(C.c_double)),
        ("n_tris", C.c_int32),
        ("tris", C.POINTER(C.c_uint32)),
        ("cloth_E", C.c_double),
        ("cloth_nu", C.c_double),
        ("cloth_bend_E", C.c_double),
        ("cloth_thickness", C.c_double),
        ("cloth_density", C.c_double),
        ("cloth_strain_rate", C.c_double),
        ("n_abd", C.c_int32),
        ("abd_body", C.POINTER(C.c_int32)),
        ("abd_kappa", C.c_double),
        ("n_faces", C.c_int32),
        ("faces", C.POINTER(C.c_uint32)),
        ("n_edges", C.c_int32),
        ("edges", C.POINTER(C.c_uint32)),
        ("n_surf_verts", C.c_int32),
        ("surf_verts", C.POINTER(C.c_uint32)),
        ("dt", C.c_double),
        ("dhat", C.c_double),
        ("gravity", C.c_double * 3),
        ("use_ground", C.c_int32),
        ("ground_normal", C.c_double * 3),
        ("ground_offset", C.c_double),
        ("kappa", C.c_double),
        ("friction_mu", C.c_double),
        ("ground_friction_mu", C.c_double),
        ("abs_xdelta_tol", C.c_double),
        ("rel_xdelta_tol", C.c_double),
        ("armijo_c1", C.c_double),
        ("armijo_beta", C.c_double),
        ("armijo_alpha_min", C.c_double),
        ("newton_max_iter", C.c_int32),
        ("pcg_rel_tol", C.c_double),
        ("pcg_abs_tol", C.c_double),
        ("pcg_max_iter", C.c_int32),
        ("pcg_use_preconditioned_norm", C.c_int32),
        ("precond", C.c_int32),
        ("mas_precision", C.c_int32),
        ("solver_precision", C.c_int32),
        ("n_soft", C.c_int32),
        ("soft_index", C.POINTER(C.c_uint32)),
        ("soft_target", C.POINTER(C.c_double)),
        ("soft_stiffness", C.c_double),
        ("motion_rate", C.c_double),
        ("bbox_diag", C.c_double),
        ("scene_diag", C.c_double),
        ("motor_speed", C.c_double),
        ("motor_strength", C.c_double),
    ]


def _ptr(a, ctype):
    if a is None:
        return C.POINTER(ctype)()
    return a.ctypes.data_as(C.POINTER(ctype))


@dataclass
class Scene:
    x: np.ndarray                    # (N,3) float64
    tets: np.ndarray                 # (T,4) uint32
    E: np.ndarray                    # (T,) float64
    nu: np.ndarray
    rho: np.ndarray
    part_id: np.ndarray              # (N,) int32 non-decreasing
    body_id: np.ndarray              # (N,) int32
    btype: np.ndarray                # (N,) int32
    faces: np.ndarray                # (F,3) uint32
    edges: np.ndarray                # (E,2) uint32
    surf_verts: np.ndarray           # (S,) uint32
    v: np.ndarray | None = None
    x_rest: np.ndarray | None = None
    tris: np.ndarray | None = None   # (M,3) uint32 cloth triangles
    cloth_E: float = 0.0
    cloth_nu: float = 0.0
    cloth_bend_E: float = 0.0
    cloth_thickness: float = 0.0
    cloth_density: float = 0.0
    cloth_strain_rate: float = 0.0
    abd_body: np.ndarray | None = None  # (N,) int32 affine body per vertex, -1 = FEM
    abd_kappa: float = 1e8
    dt: float = 0.01
    dhat: float = 1e-3
    gravity: tuple = (0.0, -9.8, 0.0)
    use_ground: bool = True
    ground_normal: tuple = (0.0, 1.0, 0.0)
    ground_offset: float = 0.0
    kappa: float = 0.0
    friction_mu: float = 0.0          # self-contact friction (reference friction_rate)
    ground_friction_mu: float = 0.0   # ground friction (reference gd_friction_rate)
    abs_xdelta_tol: float = 0.0
    rel_xdelta_tol: float = 0.0
    armijo_c1: float = 1e-4
    armijo_beta: float = 0.5
    armijo_alpha_min: float = 1e-12
    newton_max_iter: int = 10000     # reference iterCap (GIPC.cu:10739)
    pcg_rel_tol: float = 1e-4
    pcg_abs_tol: float = 0.0
    pcg_max_iter: int = 0
    pcg_use_preconditioned_norm: int = 1
    precond: int = 1
    mas_precision: int = 0
    solver_precision: int = 0
    soft_index: np.ndarray | None = None   # (K,) uint32 animated soft-target vertices
    soft_target: np.ndarray | None = None  # (K,3) float64 their targets
    soft_stiffness: float = 0.0            # motion_stiffness
    motion_rate: float = 1.0               # Newton loops per step (animation_subRate = 1/rate)
    bbox_diag: float = 0.0                 # kappa/friction scale (0: surface AABB)
    scene_diag: float = 0.0                # Newton tolerance scale (0: free-vertex AABB)
    motor_speed: float = -0.02             # motor bodies (btype 2): rad/s about body x
    motor_strength: float = 200.0          # ... and the motor's mass multiple
    name: str = "scene"
    meta: dict = field(default_factory=dict)

    @property
    def n_verts(self):
        return int(self.x.shape[0])

    @property
    def n_tets(self):
        return int(self.tets.shape[0])

    def desc(self) -> SceneDesc:
        """Build a SceneDesc whose pointers reference this object's arrays.
        Keep the Scene alive while the descriptor is used."""
        self._keep = [
            np.ascontiguousarray(self.x, np.float64),
            None if self.x_rest is None else np.ascontiguousarray(self.x_rest, np.float64),
            None if self.v is None else np.ascontiguousarray(self.v, np.float64),
            np.ascontiguousarray(self.btype, np.int32),
            np.ascontiguousarray(self.body_id, np.int32),
            np.ascontiguousarray(self.part_id, np.int32),
            np.ascontiguousarray(self.tets, np.uint32),
            np.ascontiguousarray(self.E, np.float64),
            np.ascontiguousarray(self.nu, np.float64),
            np.ascontiguousarray(self.rho, np.float64),
            np.ascontiguousarray(self.faces, np.uint32),
            np.ascontiguousarray(self.edges, np.uint32),
            np.ascontiguousarray(self.surf_verts, np.uint32),
            None if self.tris is None else np.ascontiguousarray(self.tris, np.uint32),
            None if self.abd_body is None else np.ascontiguousarray(self.abd_body, np.int32),
            None if self.soft_index is None else np.ascontiguousarray(self.soft_index, np.uint32),
            None if self.soft_target is None else np.ascontiguousarray(self.soft_target,
                                                                         np.float64),
        ]
        k = self._keep
        d = SceneDesc()
        d.n_verts = self.n_verts
        d.x = _ptr(k[0], C.c_double)
        d.x_rest = _ptr(k[1], C.c_double)
        d.v = _ptr(k[2], C.c_double)
        d.btype = _ptr(k[3], C.c_int32)
        d.body_id = _ptr(k[4], C.c_int32)
        d.part_id = _ptr(k[5], C.c_int32)
        d.n_tets = self.n_tets
        d.tets = _ptr(k[6], C.c_uint32)
        d.tet_E = _ptr(k[7], C.c_double)
        d.tet_nu = _ptr(k[8], C.c_double)
        d.tet_rho = _ptr(k[9], C.c_double)
        d.n_tris = 0 if self.tris is None else int(self.tris.shape[0])
        d.tris = _ptr(k[13], C.c_uint32)
        d.cloth_E = self.cloth_E
        d.cloth_nu = self.cloth_nu
        d.cloth_bend_E = self.cloth_bend_E
        d.cloth_thickness = self.cloth_thickness
        d.cloth_density = self.cloth_density
        d.cloth_strain_rate = self.cloth_strain_rate
        d.n_abd = 0 if self.abd_body is None else int(self.abd_body.max() + 1)
        d.abd_body = _ptr(k[14], C.c_int32)
        d.abd_kappa = self.abd_kappa
        d.n_faces = int(self.faces.shape[0])
        d.faces = _ptr(k[10], C.c_uint32)
        d.n_edges = int(self.edges.shape[0])
        d.edges = _ptr(k[11], C.c_uint32)
        d.n_surf_verts = int(self.surf_verts.shape[0])
        d.surf_verts = _ptr(k[12], C.c_uint32)
        d.dt = self.dt
        d.dhat = self.dhat
        d.gravity = (C.c_double * 3)(*self.gravity)
        d.use_ground = int(self.use_ground)
        d.ground_normal = (C.c_double * 3)(*self.ground_normal)
        d.ground_offset = self.ground_offset
        d.kappa = self.kappa
        d.friction_mu = self.friction_mu
        d.ground_friction_mu = self.ground_friction_mu
        d.abs_xdelta_tol = self.abs_xdelta_tol
        d.rel_xdelta_tol = self.rel_xdelta_tol
        d.armijo_c1 = self.armijo_c1
        d.armijo_beta = self.armijo_beta
        d.armijo_alpha_min = self.armijo_alpha_min
        d.newton_max_iter = self.newton_max_iter
        d.pcg_rel_tol = self.pcg_rel_tol
        d.pcg_abs_tol = self.pcg_abs_tol
        d.pcg_max_iter = self.pcg_max_iter
        d.pcg_use_preconditioned_norm = self.pcg_use_preconditioned_norm
        d.precond = self.precond
        d.mas_precision = self.mas_precision
        d.solver_precision = self.solver_precision
        d.n_soft = 0 if self.soft_index is None else int(np.asarray(self.soft_index).size)
        d.soft_index = _ptr(k[15], C.c_uint32)
        d.soft_target = _ptr(k[16], C.c_double)
        d.soft_stiffness = self.soft_stiffness
        d.motion_rate = self.motion_rate
        d.bbox_diag = self.bbox_diag
        d.scene_diag = self.scene_diag
        d.motor_speed = self.motor_speed
        d.motor_strength = self.motor_strength
        return d


# --------------------------------------------------------------------------
# mesh construction
# --------------------------------------------------------------------------

def hex_grid(nc, h, origin=(0.0, 0.0, 0.0), bank=(2, 2, 4)):
    """Structured grid of nc = (nx, ny, nz) hex cells of size h, each split
    into 6 tets around the main diagonal (conforming Kuhn split).

    Vertices are ordered bank by bank: the vertex lattice is cut into
    bank = (bx, by, bz) blocks (16 vertices each, smaller at the far faces);
    part_id gives the block index.  Returns (x, tets, part_id)."""
    nx, ny, nz = nc
    vx, vy, vz = nx + 1, ny + 1, nz + 1
    ii, jj, kk = np.meshgrid(np.arange(vx), np.arange(vy), np.arange(vz), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    bx, by, bz = bank
    nbx, nby, nbz = -(-vx // bx), -(-vy // by), -(-vz // bz)
    blk = (ii // bx) * (nby * nbz) + (jj // by) * nbz + (kk // bz)
    local = (ii % bx) * (by * bz) + (jj % by) * bz + (kk % bz)
    order = np.lexsort((local, blk))           # vertices sorted by (block, local)
    lattice_to_new = np.empty_like(order)
    lattice_to_new[order] = np.arange(order.size)
    blk_sorted = blk[order]
    # dense part ids in order of appearance
    _, part_id = np.unique(blk_sorted, return_inverse=True)
    x = np.stack([ii, jj, kk], 1)[order].astype(np.float64) * h + np.asarray(origin, np.float64)

    def vid(i, j, k):
        return lattice_to_new[(i * vy + j) * vz + k]

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    e = np.eye(3, dtype=np.int64)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        a = np.zeros(3, np.int64)
        b = e[p[0]]
        c = e[p[0]] + e[p[1]]
        d = np.ones(3, np.int64)
        tets.append(np.stack([vid(ci + q[0], cj + q[1], ck + q[2]) for q in (a, b, c, d)], 1))
    tets = np.concatenate(tets, 0).astype(np.uint32)
    return x, tets, part_id.astype(np.int32)


def surface_of(tets):
    """Boundary triangles (faces seen once), their unique edges and vertices."""
    t = np.asarray(tets, np.int64)
    f = np.concatenate([t[:, [0, 1, 2]], t[:, [0, 1, 3]], t[:, [0, 2, 3]], t[:, [1, 2, 3]]], 0)
    key = np.sort(f, 1)
    _, idx, cnt = np.unique(key, axis=0, return_index=True, return_counts=True)
    faces = f[idx[cnt == 1]]
    faces = faces[np.lexsort((faces[:, 2], faces[:, 1], faces[:, 0]))]
    e = np.concatenate([faces[:, [0, 1]], faces[:, [1, 2]], faces[:, [2, 0]]], 0)
    e = np.unique(np.sort(e, 1), axis=0)
    sv = np.unique(faces.ravel())
    return faces.astype(np.uint32), e.astype(np.uint32), sv.astype(np.uint32)


def cloth_grid(n, size, origin=(0.0, 0.0, 0.0), bank=(4, 4), normal_axis=1):
    """Square sheet of n x n vertices spanning size metres in the plane normal
    to normal_axis, 2 (n-1)^2 triangles (alternating diagonals).  Vertices
    are ordered bank by bank (4x4 blocks = one 16-vertex MAS bank, part ids
    non-decreasing).  Returns (x, tris, part_id)."""
    ii, jj = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.ravel(), jj.ravel()
    bi, bj = bank
    nbj = -(-n // bj)
    blk = (ii // bi) * nbj + (jj // bj)
    local = (ii % bi) * bj + (jj % bj)
    order = np.lexsort((local, blk))
    lattice_to_new = np.empty_like(order)
    lattice_to_new[order] = np.arange(order.size)
    _, part_id = np.unique(blk[order], return_inverse=True)
    h = size / (n - 1)
    a, b = [k for k in range(3) if k != normal_axis]
    x = np.zeros((n * n, 3))
    x[:, a] = ii[order] * h
    x[:, b] = jj[order] * h
    x += np.asarray(origin, np.float64)

    def vid(i, j):
        return lattice_to_new[i * n + j]

    ci, cj = np.meshgrid(np.arange(n - 1), np.arange(n - 1), indexing="ij")
    ci, cj = ci.ravel(), cj.ravel()
    v00, v10, v01, v11 = vid(ci, cj), vid(ci + 1, cj), vid(ci, cj + 1), vid(ci + 1, cj + 1)
    flip = (ci + cj) % 2 == 1
    t1 = np.where(flip[:, None], np.stack([v00, v10, v01], 1), np.stack([v00, v10, v11], 1))
    t2 = np.where(flip[:, None], np.stack([v10, v11, v01], 1), np.stack([v00, v11, v01], 1))
    tris = np.concatenate([t1, t2], 0).astype(np.uint32)
    return x, tris, part_id.astype(np.int32)


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def assemble_bodies(bodies, cloths=(), **kw):
    """Concatenate bodies = [(x, tets, part_id, E, nu, rho, btype[, abd])] and
    cloths = [(x, tris, part_id, btype)] into a Scene with per-body ids and
    offset part ids.  btype is a scalar or a per-vertex array.  Bodies with
    abd = True become affine bodies (their tets supply mass and volume only)
    and are placed after every FEM and cloth vertex.  Cloth vertices get body
    id -1 (self contact allowed); the collision surface is the tet boundary
    plus every cloth triangle, edge and vertex."""
    xs, ts, ps, Es, nus, rhos, bts, bids, trs = [], [], [], [], [], [], [], [], []
    voff = 0
    poff = 0
    fem = [(i, b) for i, b in enumerate(bodies) if not (len(b) > 7 and b[7])]
    abd = [(i, b) for i, b in enumerate(bodies) if len(b) > 7 and b[7]]
    abd_of = []
    for b, (x, t, p, E, nu, rho, bt) in [(i, bb[:7]) for i, bb in fem]:
        abd_of.append(np.full(x.shape[0], -1, np.int32))
        xs.append(x)
        ts.append(t.astype(np.int64) + voff)
        ps.append(p + poff)
        Es.append(np.full(t.shape[0], E))
        nus.append(np.full(t.shape[0], nu))
        rhos.append(np.full(t.shape[0], rho))
        bts.append(np.broadcast_to(np.asarray(bt, np.int32), (x.shape[0],)).copy())
        bids.append(np.full(x.shape[0], b, np.int32))
        voff += x.shape[0]
        poff = int(ps[-1].max()) + 1
    for (x, tr, p, bt) in cloths:
        xs.append(x)
        trs.append(tr.astype(np.int64) + voff)
        ps.append(p + poff)
        bts.append(np.broadcast_to(np.asarray(bt, np.int32), (x.shape[0],)).copy())
        bids.append(np.full(x.shape[0], -1, np.int32))
        abd_of.append(np.full(x.shape[0], -1, np.int32))
        voff += x.shape[0]
        poff = int(ps[-1].max()) + 1
    for k, (b, (x, t, p, E, nu, rho, bt)) in enumerate([(i, bb[:7]) for i, bb in abd]):
        xs.append(x)
        ts.append(t.astype(np.int64) + voff)
        ps.append(p + poff)
        Es.append(np.full(t.shape[0], E))
        nus.append(np.full(t.shape[0], nu))
        rhos.append(np.full(t.shape[0], rho))
        bts.append(np.broadcast_to(np.asarray(bt, np.int32), (x.shape[0],)).copy())
        bids.append(np.full(x.shape[0], b, np.int32))
        abd_of.append(np.full(x.shape[0], k, np.int32))
        voff += x.shape[0]
        poff = int(ps[-1].max()) + 1
    tets = np.concatenate(ts).astype(np.uint32) if ts else np.zeros((0, 4), np.uint32)
    faces, edges, sv = surface_of(tets)
    tris = None
    if trs:
        tris = np.concatenate(trs).astype(np.uint32)
        te = np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]], 0)
        faces = np.concatenate([faces, tris], 0).astype(np.uint32)
        edges = np.unique(np.concatenate([edges.astype(np.int64), np.sort(te, 1)], 0),
                          axis=0).astype(np.uint32)
        sv = np.unique(np.concatenate([sv, tris.ravel()])).astype(np.uint32)
    cat = (lambda a: np.concatenate(a) if a else np.zeros(0))
    return Scene(
        x=np.concatenate(xs), tets=tets, E=cat(Es), nu=cat(nus),
        rho=cat(rhos), part_id=np.concatenate(ps).astype(np.int32),
        body_id=np.concatenate(bids), btype=np.concatenate(bts), faces=faces, edges=edges,
        surf_verts=sv, tris=tris, abd_body=np.concatenate(abd_of) if abd else None, **kw)


# --------------------------------------------------------------------------
# the benchmark scenes (sizes of the metric's configurations)
# --------------------------------------------------------------------------

def with_soft_targets(sc, index, targets=None, stiffness=1e3, motion_rate=1.0):
    """A copy of the scene whose vertices `index` are animated soft targets
    (reference scene JSON "animation" constraints, settings motion_stiffness /
    motion_rate), pulled towards `targets` (default: their positions)."""
    idx = np.ascontiguousarray(index, np.uint32)
    t = sc.x[idx].copy() if targets is None else np.array(targets, np.float64).reshape(-1, 3)
    return replace(sc, soft_index=idx, soft_target=t, soft_stiffness=float(stiffness),
                   motion_rate=float(motion_rate))


def rotate_axis_angle(p, origin, axis, angle):
    """Rotate points p about the axis through origin by angle (Rodrigues; the
    frame update of animated targets, main.cu rotate_axis_angle)."""
    k = np.asarray(axis, np.float64)
    k = k / np.linalg.norm(k)
    q = np.asarray(p, np.float64) - origin
    c, s = np.cos(angle), np.sin(angle)
    return origin + q * c + np.cross(k, q) * s + np.outer(q @ k, k) * (1 - c)


def bar_scene(cells=(16, 4, 4), h=0.05, drop=0.1, E=1e5, nu=0.49, rho=1000.0, **kw):
    """One tet bar of 16x4x4 hex cells (h = 0.05 m, 425 vertices) dropped from
    0.1 m onto the ground plane; SNK1 E = 1e5, nu = 0.49, rho = 1000."""
    x, t, p = hex_grid(cells, h, origin=(0.0, drop, 0.0))
    defaults = dict(dt=0.01, dhat=1e-3, name="bar_16x4x4")
    defaults.update(kw)
    sc = assemble_bodies([(x, t, p, E, nu, rho, 0)], **defaults)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc


def cubes_scene(n_side=4, cells=8, size=0.2, spacing=0.6, fall=0.5, seed=0, nu=0.4,
                rho=1000.0, E_range=(1e5, 1e9), crop=None, **kw):
    """n_side^3 tet cubes of cells^3 hex cells placed on a grid with seeded
    +/-15% jitter and uniform random rotations, Young's modulus log-uniform in
    E_range, lowest cube 'fall' metres above the ground.  crop = (na, nb, nc)
    keeps only the cubes at grid index (a < na, b < nb, c < nc) (b vertical),
    each with the same position, rotation and modulus as in the full scene."""
    rng = np.random.default_rng(seed)
    x0, t0, p0 = hex_grid((cells,) * 3, size / cells)
    x0 = x0 - x0.mean(0)
    half_diag = np.sqrt(3.0) * size / 2
    bodies = []
    for a in range(n_side):
        for b in range(n_side):
            for c in range(n_side):
                R = random_rotation(rng)
                jit = rng.uniform(-0.15, 0.15, 3) * spacing
                ctr = np.array([a * spacing, fall + half_diag + b * spacing, c * spacing]) + jit
                ctr[1] = max(ctr[1], fall + half_diag)
                E = float(np.exp(rng.uniform(np.log(E_range[0]), np.log(E_range[1]))))
                if crop is not None and not (a < crop[0] and b < crop[1] and c < crop[2]):
                    continue
                bodies.append((x0 @ R.T + ctr, t0, p0, E, nu, rho, 0))
    defaults = dict(dt=0.01, dhat=1e-3, name=f"cubes_{len(bodies)}x{cells}^3")
    defaults.update(kw)
    sc = assemble_bodies(bodies, **defaults)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc


def _diag(x):
    return float(np.linalg.norm(x.max(0) - x.min(0)))


def stack_scene(n=2, cells=3, size=0.1, gap=0.5e-3, rot_deg=7.0, seed=0, E=1e6, nu=0.4,
                rho=1000.0, **kw):
    """n small tet cubes stacked on the ground, each rotated by rot_deg about
    the vertical axis (alternating sign) with a gap of 'gap' metres between
    neighbours and to the ground: contacts (ground, PT, EE) from step 0."""
    x0, t0, p0 = hex_grid((cells,) * 3, size / cells)
    x0 = x0 - x0.mean(0)
    bodies = []
    y = gap + size / 2
    for k in range(n):
        a = np.deg2rad(rot_deg * (1 if k % 2 == 0 else -1) * (k > 0))
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        bodies.append((x0 @ R.T + np.array([0.0, y, 0.0]), t0, p0, E, nu, rho, 0))
        y += size + gap
    defaults = dict(dt=0.01, dhat=1e-3, name=f"stack{n}x{cells}^3")
    defaults.update(kw)
    sc = assemble_bodies(bodies, **defaults)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc


def cloth_scene(n=16, size=0.3, height=0.15, n_boxes=1, box_size=0.1, box_cells=2, seed=0,
                E=1e6, nu=0.3, bend_stiffness=1e-4, thickness=1e-3, density=200.0,
                strain_rate=0.1, box_E=1e9, box_rho=1000.0, fixed_boxes=True, abd_boxes=False,
                **kw):
    """Cloth sheet of n x n vertices (size metres square, 2 (n-1)^2 triangles)
    held horizontally at 'height' above n_boxes tet boxes of box_size metres
    resting on the ground at seeded random positions under the sheet (one
    box: centred).  Membrane E, nu; the bending stiffness k = E_b h^3 /
    (24 (1 - nu^2)) is set directly through E_b = 24 (1 - nu^2) k / h^3.
    Boxes are fixed obstacles (btype 1) unless fixed_boxes is False; with
    abd_boxes they are free affine bodies (kappa from abd_kappa)."""
    rng = np.random.default_rng(seed)
    xc, tc, pc = cloth_grid(n, size, origin=(-size / 2, height, -size / 2))
    xb0, tb0, pb0 = hex_grid((box_cells,) * 3, box_size / box_cells)
    xb0 = xb0 - xb0.mean(0)
    bodies = []
    placed = []
    for k in range(n_boxes):  # n_boxes = 0: the sheet and the ground only
        if n_boxes == 1:
            ctr = np.array([0.0, 0.0, 0.0])
        else:
            # seeded positions, redrawn until the (axis-aligned) boxes are at
            # least 2 cm apart
            for _ in range(10000):
                ctr = np.array([rng.uniform(-0.4, 0.4) * size, 0.0,
                                rng.uniform(-0.4, 0.4) * size])
                if all(max(abs(ctr[0] - p[0]), abs(ctr[2] - p[2])) >= box_size + 0.02
                       for p in placed):
                    break
            else:
                raise ValueError("cloth_scene: no room for the boxes under the sheet")
            placed.append(ctr)
        ctr[1] = box_size / 2 + 1e-3 * (1 + k % 3)
        if abd_boxes:
            bodies.append((xb0 + ctr, tb0, pb0, box_E, 0.3, box_rho, 0, True))
        else:
            bodies.append((xb0 + ctr, tb0, pb0, box_E, 0.3, box_rho, 1 if fixed_boxes else 0))
    bend_E = 24.0 * (1.0 - nu * nu) * bend_stiffness / thickness ** 3
    defaults = dict(dt=0.01, dhat=1e-3, name=f"cloth{n}x{n}_{n_boxes}boxes", cloth_E=E,
                    cloth_nu=nu, cloth_bend_E=bend_E, cloth_thickness=thickness,
                    cloth_density=density, cloth_strain_rate=strain_rate)
    defaults.update(kw)
    sc = assemble_bodies(bodies, cloths=[(xc, tc, pc, 0)], **defaults)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc


def abd_scene(n_abd=2, cells=2, size=0.1, gap=0.5e-3, fem_base=True, base_E=1e6, rot_deg=9.0,
              abd_kappa=1e8, rho=1000.0, seed=0, **kw):
    """A soft FEM box on the ground (fem_base) with n_abd affine boxes
    stacked on it, each rotated by rot_deg about the vertical axis
    (alternating) and 'gap' above its support: contacts between affine
    bodies, between an affine body and the FEM box, and with the ground."""
    x0, t0, p0 = hex_grid((cells,) * 3, size / cells)
    x0 = x0 - x0.mean(0)
    bodies = []
    y = gap + size / 2
    if fem_base:
        bodies.append((x0 + np.array([0.0, y, 0.0]), t0, p0, base_E, 0.4, rho, 0))
        y += size + gap
    for k in range(n_abd):
        a = np.deg2rad(rot_deg * (1 if k % 2 == 0 else -1))
        R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        bodies.append((x0 @ R.T + np.array([0.0, y, 0.0]), t0, p0, 1e9, 0.3, rho, 0, True))
        y += size + gap
    defaults = dict(dt=0.01, dhat=1e-3, name=f"abd{n_abd}", abd_kappa=abd_kappa)
    defaults.update(kw)
    sc = assemble_bodies(bodies, **defaults)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc


def hybrid_scene(n_side=8, box_cells=6, bed_cells=(64, 40, 64), bed_size=1.6, box_size=0.12,
                 speed=10.0, gap=5e-3, bed_E=5e4, nu=0.4, rho=1000.0, abd_kappa=1e11, mu=0.3,
                 max_tilt_deg=15.0, seed=0, **kw):
    """Stiff/soft impact: n_side^3 affine boxes of box_cells^3 hex cells
    (box_size metres) on an n_side^3 lattice over a soft tet bed of
    bed_cells = (nx, ny, nz) hex cells (ny vertical, bed_size metres along x
    and z, bottom layer fixed), all boxes moving down at 'speed' m/s, the
    lowest layer 'gap' above the bed.  Boxes are tilted by seeded rotations
    of at most max_tilt_deg; friction mu for self contact.  Defaults: 512
    boxes of 1,296 tets into a 983,040-tet bed, dt = 0.005, dhat = 5e-4."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = bed_cells
    h = bed_size / nx
    xb, tb, pb = hex_grid(bed_cells, h, origin=(-bed_size / 2, 0.0, -bed_size / 2 * nz / nx))
    bt_bed = (xb[:, 1] <= 1e-12).astype(np.int32)      # bottom layer fixed
    top = ny * h
    x0, t0, p0 = hex_grid((box_cells,) * 3, box_size / box_cells)
    x0 = x0 - x0.mean(0)
    pitch = bed_size / n_side
    bodies = [(xb, tb, pb, bed_E, nu, rho, bt_bed)]
    vel = [np.zeros_like(xb)]
    for a in range(n_side):
        for b in range(n_side):
            for c in range(n_side):
                ax = rng.normal(size=3)
                ax /= np.linalg.norm(ax)
                ang = np.deg2rad(rng.uniform(-max_tilt_deg, max_tilt_deg))
                K = np.array([[0, -ax[2], ax[1]], [ax[2], 0, -ax[0]], [-ax[1], ax[0], 0]])
                R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
                xr = x0 @ R.T
                ctr = np.array([(a + 0.5) * pitch - bed_size / 2, 0.0,
                                (c + 0.5) * pitch - bed_size / 2 * nz / nx])
                ctr[1] = top + gap - xr[:, 1].min() + b * pitch
                bodies.append((xr + ctr, t0, p0, 1e9, 0.3, rho, 0, True))
                vel.append(np.tile([0.0, -speed, 0.0], (xr.shape[0], 1)))
    defaults = dict(dt=0.005, dhat=5e-4, name=f"hybrid_{n_side ** 3}abd_bed{nx}x{ny}x{nz}",
                    abd_kappa=abd_kappa, friction_mu=mu, ground_friction_mu=mu,
                    use_ground=False)
    defaults.update(kw)
    sc = assemble_bodies(bodies, **defaults)
    # assemble_bodies places FEM vertices first and affine bodies after, in
    # body order, which is the order of `vel`
    sc.v = np.concatenate(vel)
    sc.abs_xdelta_tol = 1e-2 * sc.dt / _diag(sc.x)
    return sc
