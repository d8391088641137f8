"""ctypes bindings of oracle/liboracle.so -- the CPU FP64 checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  The product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
up = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


class Bsr(C.Structure):
    _fields_ = [("nb", C.c_int), ("nnzb", C.c_int), ("rowptr", C.POINTER(C.c_int)),
                ("cols", C.POINTER(C.c_int)), ("vals", dp), ("nbody", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_sym_eig.argtypes = [C.c_int, dp, dp, dp]
        L.orc_psd_project.argtypes = [C.c_int, dp]
        L.orc_inverse.argtypes = [C.c_int, dp]
        L.orc_snk1_psi.restype = C.c_double
        L.orc_snk1_psi.argtypes = [dp, C.c_double, C.c_double]
        L.orc_snk1_pk1.argtypes = [dp, C.c_double, C.c_double, dp]
        L.orc_snk1_hess.argtypes = [dp, C.c_double, C.c_double, C.c_int, dp]
        L.orc_fem_precompute.argtypes = [C.c_int, dp, C.c_int, up, dp, dp, dp, dp, dp, dp, dp, dp]
        L.orc_tet_grad_hess.argtypes = [dp, up, dp, C.c_double, C.c_double, C.c_double,
                                        C.c_double, dp, dp]
        L.orc_tet_energy.restype = C.c_double
        L.orc_tet_energy.argtypes = [dp, up, dp, C.c_double, C.c_double, C.c_double]
        L.orc_ground_energy.restype = C.c_double
        L.orc_ground_energy.argtypes = [C.c_double, C.c_double]
        L.orc_ground_grad_hess.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double, dp, dp]
        L.orc_expand_bits.restype = C.c_uint32
        L.orc_expand_bits.argtypes = [C.c_uint32]
        L.orc_morton_code.restype = C.c_uint32
        L.orc_morton_code.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_hash_code.restype = C.c_uint32
        L.orc_hash_code.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_tet_hash_type.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_spatial_keys.argtypes = [C.c_int, dp, C.c_int, up, C.c_int, u64p, up]
        L.orc_leaf_aabbs.argtypes = [dp, up, C.c_int, C.c_int, dp]
        L.orc_lbvh_build.argtypes = [dp, up, C.c_int, C.c_int, u64p, up, dp]
        for f in ("orc_d_pp",):
            getattr(L, f).restype = C.c_double
        L.orc_d_pp.argtypes = [dp, dp]
        L.orc_d_pe.restype = C.c_double
        L.orc_d_pe.argtypes = [dp, dp, dp]
        L.orc_d_pt.restype = C.c_double
        L.orc_d_pt.argtypes = [dp, dp, dp, dp]
        L.orc_d_ee.restype = C.c_double
        L.orc_d_ee.argtypes = [dp, dp, dp, dp]
        L.orc_dtype_pt.argtypes = [dp, dp, dp, dp]
        L.orc_dtype_ee.argtypes = [dp, dp, dp, dp]
        L.orc_contact_pairs_bruteforce.argtypes = [C.c_void_p, dp, C.c_double, ip, C.c_int, ip]
        L.orc_contact_pairs_bvh.argtypes = [C.c_void_p, dp, C.c_double, ip, C.c_int, ip]
        L.orc_barrier_energy.restype = C.c_double
        L.orc_barrier_energy.argtypes = [dp, dp, ip, C.c_double, C.c_double]
        L.orc_barrier_grad_hess.argtypes = [dp, dp, ip, C.c_double, C.c_double, ip, dp, dp]
        L.orc_barrier_lambda0.restype = C.c_double
        L.orc_barrier_lambda0.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_bsr_from_upper_triplets.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int),
                                                  C.POINTER(C.c_int), dp, C.POINTER(Bsr)]
        L.orc_bsr_free.argtypes = [C.POINTER(Bsr)]
        L.orc_bsr_spmv.argtypes = [C.POINTER(Bsr), dp, dp]
        L.orc_mas_create.restype = C.c_void_p
        L.orc_mas_create.argtypes = [C.POINTER(Bsr), ip, C.c_int]
        L.orc_mas_free.argtypes = [C.c_void_p]
        L.orc_mas_apply.argtypes = [C.c_void_p, dp, dp]
        L.orc_mas_levels.argtypes = [C.c_void_p]
        L.orc_mas_level_sizes.argtypes = [C.c_void_p, ip]
        L.orc_mas_total_clusters.argtypes = [C.c_void_p]
        L.orc_mas_chain.argtypes = [C.c_void_p, ip]
        L.orc_block_jacobi.argtypes = [C.POINTER(Bsr), dp, dp]
        L.orc_abd_motor_target.argtypes = [dp, C.c_double, dp]
        L.orc_pcg.argtypes = [C.POINTER(Bsr), C.c_void_p, C.c_int, dp, dp, C.c_double,
                              C.c_double, C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_world_create.restype = C.c_void_p
        L.orc_world_create.argtypes = [C.c_void_p]
        L.orc_world_free.argtypes = [C.c_void_p]
        L.orc_world_set_positions.argtypes = [C.c_void_p, dp]
        L.orc_world_set_positions_exact.argtypes = [C.c_void_p, dp]
        L.orc_world_set_soft_targets.argtypes = [C.c_void_p, dp]
        L.orc_world_set_anim_rate.argtypes = [C.c_void_p, C.c_double]
        L.orc_world_get_anim_rate.restype = C.c_double
        L.orc_world_get_anim_rate.argtypes = [C.c_void_p]
        L.orc_world_get_positions.argtypes = [C.c_void_p, dp]
        L.orc_world_set_xtilde.argtypes = [C.c_void_p, dp]
        L.orc_world_get_xtilde.argtypes = [C.c_void_p, dp]
        L.orc_world_get_kappa.restype = C.c_double
        L.orc_world_get_kappa.argtypes = [C.c_void_p]
        L.orc_world_set_kappa.argtypes = [C.c_void_p, C.c_double]
        L.orc_world_assemble.argtypes = [C.c_void_p, dp, C.POINTER(Bsr)]
        L.orc_world_energy.argtypes = [C.c_void_p, dp]
        L.orc_world_newton_iteration.argtypes = [C.c_void_p, dp, ip, dp, ip]
        L.orc_world_time_step.argtypes = [C.c_void_p, ip, ip]
        L.orc_world_set_state.argtypes = [C.c_void_p, dp, dp, dp, C.c_double]
        L.orc_world_advance.argtypes = [C.c_void_p, C.c_int, ip, ip]
        L.orc_friction_lag.argtypes = [dp, ip, C.c_double, C.c_double, dp, dp, dp]
        L.orc_friction_lag_ground.restype = C.c_double
        L.orc_friction_lag_ground.argtypes = [dp, dp, C.c_double, C.c_double, C.c_double]
        L.orc_friction_energy.restype = C.c_double
        L.orc_friction_energy.argtypes = [dp, dp, ip, dp, dp, C.c_double, C.c_double]
        L.orc_friction_grad_hess.argtypes = [dp, dp, ip, dp, dp, C.c_double, C.c_double,
                                             C.c_double, ip, dp, dp]
        L.orc_friction_ground_energy.restype = C.c_double
        L.orc_friction_ground_energy.argtypes = [dp, dp, dp, C.c_double, C.c_double]
        L.orc_friction_ground_grad_hess.argtypes = [dp, dp, dp, C.c_double, C.c_double,
                                                    C.c_double, dp, dp]
        d = C.c_double
        L.orc_cloth_params.argtypes = [d, d, d, d, d, dp, dp, dp]
        L.orc_tri_precompute.argtypes = [dp, up, dp, dp]
        L.orc_bw_psi.restype = d
        L.orc_bw_psi.argtypes = [dp, d, d, d]
        L.orc_bw_pk1.argtypes = [dp, d, d, d, dp]
        L.orc_bw_hess.argtypes = [dp, d, d, d, C.c_int, dp]
        L.orc_tri_energy.restype = d
        L.orc_tri_energy.argtypes = [dp, up, dp, d, d, d, d]
        L.orc_tri_grad_hess.argtypes = [dp, up, dp, d, d, d, d, d, dp, dp]
        L.orc_dihedral.restype = d
        L.orc_dihedral.argtypes = [dp, dp, dp, dp, dp, dp]
        L.orc_bend_energy.restype = d
        L.orc_bend_energy.argtypes = [dp, up, d, d, d]
        L.orc_bend_grad_hess.argtypes = [dp, up, d, d, d, d, dp, dp]
        L.orc_bend_hinges.restype = C.c_int
        L.orc_bend_hinges.argtypes = [C.c_int, up, up]
        L.orc_abd_shape_energy.restype = d
        L.orc_abd_shape_energy.argtypes = [dp]
        L.orc_abd_shape_grad.argtypes = [dp, dp]
        L.orc_abd_shape_hess.argtypes = [dp, dp]
        L.orc_abd_x.argtypes = [dp, dp, dp]
        L.orc_abd_jt.argtypes = [dp, dp, dp]
        L.orc_abd_jmat.argtypes = [dp, dp]
        L.orc_bf16_round.restype = C.c_float
        L.orc_bf16_round.argtypes = [C.c_float]
        L.orc_mlp_params.restype = C.c_int
        L.orc_mlp_loss_grad.restype = d
        L.orc_mlp_loss_grad.argtypes = [C.POINTER(C.c_float)] * 3 + [C.c_int, dp,
                                                                      C.POINTER(C.c_float)]
        L.orc_world_nodes.restype = C.c_int
        L.orc_world_nodes.argtypes = [C.c_void_p]
        L.orc_pt_ccd.restype = C.c_double
        L.orc_pt_ccd.argtypes = [dp] * 8 + [C.c_double, C.c_double]
        L.orc_ee_ccd.restype = C.c_double
        L.orc_ee_ccd.argtypes = [dp] * 8 + [C.c_double, C.c_double]
        L.orc_inverse_gj.argtypes = [C.c_int, dp]
        L.orc_world_ccd_bound.argtypes = [C.c_void_p, dp, dp]
        L.orc_world_ccd_candidates.restype = C.c_int
        L.orc_world_ccd_candidates.argtypes = [C.c_void_p, dp, C.c_double, ip, C.c_int]
        L.orc_barrier_H_b.restype = d
        L.orc_barrier_H_b.argtypes = [d, d]
        L.orc_world_kappa_info.argtypes = [C.c_void_p, dp]
        L.orc_world_init_kappa.restype = d
        L.orc_world_init_kappa.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_get_threads.restype = C.c_int
        _lib = L
    return _lib


def set_threads(n):
    """Host threads of the oracle's parallel loops (1: sequential).  Every
    parallel sum uses fixed chunks combined in index order, so results do not
    depend on the thread count."""
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


def inverse_gj(M):
    """In-place Gauss-Jordan inverse without pivoting in the reference's
    operation order (MASPreconditioner.cu:628-724); returns a new array."""
    A = np.ascontiguousarray(M, np.float64).copy()
    lib().orc_inverse_gj(A.shape[0], P(A))
    return A


def barrier_H_b(d, dhat):
    return lib().orc_barrier_H_b(float(d), float(dhat))


def P(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


# ----------------------------------------------------------------- helpers

def snk1(F, mu, lam):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    Pm = np.zeros(9)
    H = np.zeros(81)
    lib().orc_snk1_pk1(P(F), mu, lam, P(Pm))
    return lib().orc_snk1_psi(P(F), mu, lam), Pm.reshape(3, 3)


def snk1_hess(F, mu, lam, project=True):
    F = np.ascontiguousarray(F, np.float64).reshape(9)
    H = np.zeros(81)
    lib().orc_snk1_hess(P(F), mu, lam, int(project), P(H))
    return H.reshape(9, 9)


def lbvh(x, prims):
    x = np.ascontiguousarray(x, np.float64)
    prims = np.ascontiguousarray(prims, np.uint32)
    n, ar = prims.shape
    keys = np.zeros(n, np.uint64)
    nodes = np.zeros((2 * n - 1, 4), np.uint32)
    aabbs = np.zeros((2 * n - 1, 6), np.float64)
    lib().orc_lbvh_build(P(x), P(prims, C.c_uint32), ar, n, P(keys, C.c_uint64),
                         P(nodes, C.c_uint32), P(aabbs))
    return keys, nodes, aabbs


def spatial_keys(x, tets=None):
    """sorted keys (code << 32 | index) and the order of the vertices
    (tets None) or of the tets (GIPC.cu:279-461)"""
    x = np.ascontiguousarray(x, np.float64)
    nv = x.shape[0]
    if tets is None:
        kind, n, t = 0, nv, np.zeros((1, 4), np.uint32)
    else:
        t = np.ascontiguousarray(tets, np.uint32)
        kind, n = 1, t.shape[0]
    keys = np.zeros(n, np.uint64)
    order = np.zeros(n, np.uint32)
    r = lib().orc_spatial_keys(kind, P(x), nv, P(t, C.c_uint32), t.shape[0] if kind else 0,
                               P(keys, C.c_uint64), P(order, C.c_uint32))
    if r != 0:
        raise ValueError("no vertices or no tets")
    return keys, order


def contact_pairs(scene, x=None, method="bvh", cap=None):
    d = scene.desc()
    x = np.ascontiguousarray(scene.x if x is None else x, np.float64)
    cap = cap or max(1000, 4 * (scene.surf_verts.size + scene.edges.shape[0]))
    pairs = np.zeros((cap, 4), np.int32)
    counts = np.zeros(5, np.int32)
    f = lib().orc_contact_pairs_bvh if method == "bvh" else lib().orc_contact_pairs_bruteforce
    n = f(C.byref(d), P(x), scene.dhat ** 2, P(pairs, C.c_int32), cap, P(counts, C.c_int32))
    assert n <= cap, "pair capacity exceeded"
    return pairs[:n].copy(), counts


class OBsr:
    """Owned full BSR produced by the oracle."""

    def __init__(self, raw: Bsr):
        self.nb = raw.nb
        self.nnzb = raw.nnzb
        self.rowptr = np.ctypeslib.as_array(raw.rowptr, (raw.nb + 1,)).copy()
        self.cols = np.ctypeslib.as_array(raw.cols, (max(raw.nnzb, 1),))[: raw.nnzb].copy()
        self.vals = np.ctypeslib.as_array(raw.vals, (max(raw.nnzb, 1) * 9,))[: 9 * raw.nnzb] \
            .reshape(raw.nnzb, 3, 3).copy()
        self.nbody = raw.nbody
        lib().orc_bsr_free(C.byref(raw))

    def raw(self):
        self._keep = (np.ascontiguousarray(self.rowptr, np.int32),
                      np.ascontiguousarray(self.cols, np.int32),
                      np.ascontiguousarray(self.vals, np.float64))
        r = Bsr()
        r.nb, r.nnzb = self.nb, self.nnzb
        r.rowptr = P(self._keep[0], C.c_int)
        r.cols = P(self._keep[1], C.c_int)
        r.vals = P(self._keep[2])
        r.nbody = getattr(self, "nbody", 0)
        return r

    def dense(self):
        n = 3 * self.nb
        A = np.zeros((n, n))
        for i in range(self.nb):
            for k in range(self.rowptr[i], self.rowptr[i + 1]):
                j = self.cols[k]
                A[3 * i:3 * i + 3, 3 * j:3 * j + 3] = self.vals[k]
        return A

    def spmv(self, x):
        y = np.zeros(3 * self.nb)
        r = self.raw()
        lib().orc_bsr_spmv(C.byref(r), P(np.ascontiguousarray(x, np.float64)), P(y))
        return y


def bsr_from_upper_triplets(nb, rows, cols, vals):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    raw = Bsr()
    rc = lib().orc_bsr_from_upper_triplets(nb, rows.size, P(rows, C.c_int), P(cols, C.c_int),
                                           P(vals), C.byref(raw))
    assert rc == 0
    return OBsr(raw)


class Mas:
    def __init__(self, A: OBsr, part_id=None, fp32=True):
        self._raw = A.raw()
        self._part = None if part_id is None else np.ascontiguousarray(part_id, np.int32)
        self.h = lib().orc_mas_create(C.byref(self._raw),
                                      None if self._part is None else P(self._part, C.c_int32),
                                      int(fp32))
        self.nb = A.nb

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mas_free(self.h)
            self.h = None

    def apply(self, r):
        z = np.zeros(3 * self.nb)
        lib().orc_mas_apply(self.h, P(np.ascontiguousarray(r, np.float64)), P(z))
        return z

    @property
    def levels(self):
        return lib().orc_mas_levels(self.h)

    def level_sizes(self):
        s = np.zeros(2 * (self.levels + 1), np.int32)
        lib().orc_mas_level_sizes(self.h, P(s, C.c_int32))
        return s.reshape(-1, 2)

    def total_clusters(self):
        return lib().orc_mas_total_clusters(self.h)

    def chain(self):
        L = self.levels
        c = np.zeros(self.nb * max(L - 1, 1), np.int32)
        lib().orc_mas_chain(self.h, P(c, C.c_int32))
        return c.reshape(self.nb, max(L - 1, 1))


def abd_motor_target(qt, theta):
    """motor target q_p of an affine body from its prediction q~"""
    q = np.ascontiguousarray(qt, np.float64)
    out = np.zeros(12)
    lib().orc_abd_motor_target(P(q), float(theta), P(out))
    return out


def block_jacobi(A: OBsr, r):
    """z = block-Jacobi(A) r: 3x3 node blocks, each affine body (the last
    4 nbody nodes) by its 12x12 diagonal block"""
    raw = A.raw()
    z = np.zeros(3 * A.nb)
    lib().orc_block_jacobi(C.byref(raw), P(np.ascontiguousarray(r, np.float64)), P(z))
    return z


def pcg(A: OBsr, b, mas: Mas | None = None, precond=1, rel_tol=1e-4, abs_tol=0.0, max_iter=0,
        use_rz=1, check_every=10):
    raw = A.raw()
    x = np.zeros(3 * A.nb)
    rz0 = C.c_double()
    rz = C.c_double()
    it = lib().orc_pcg(C.byref(raw), mas.h if mas else None, precond,
                       P(np.ascontiguousarray(b, np.float64)), P(x), rel_tol, abs_tol, max_iter,
                       use_rz, check_every, C.byref(rz0), C.byref(rz))
    return x, it, rz0.value, rz.value


class World:
    """Oracle simulation state for one scene."""

    def __init__(self, scene):
        self.scene = scene
        self._desc = scene.desc()
        self.h = lib().orc_world_create(C.byref(self._desc))
        self.nv = scene.n_verts

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_world_free(self.h)
            self.h = None

    def positions(self):
        x = np.zeros(3 * self.nv)
        lib().orc_world_get_positions(self.h, P(x))
        return x.reshape(-1, 3)

    def set_positions(self, x):
        lib().orc_world_set_positions(self.h, P(np.ascontiguousarray(x, np.float64).ravel()))

    def set_positions_exact(self, x):
        """vertex positions kept exactly as given (affine bodies refitted,
        their vertices not projected): position-level comparisons"""
        lib().orc_world_set_positions_exact(self.h, P(np.ascontiguousarray(x, np.float64).ravel()))

    def xtilde(self):
        x = np.zeros(3 * self.nv)
        lib().orc_world_get_xtilde(self.h, P(x))
        return x.reshape(-1, 3)

    def set_xtilde(self, xt):
        lib().orc_world_set_xtilde(self.h, P(np.ascontiguousarray(xt, np.float64).ravel()))

    @property
    def kappa(self):
        return lib().orc_world_get_kappa(self.h)

    @kappa.setter
    def kappa(self, k):
        lib().orc_world_set_kappa(self.h, k)

    @property
    def nodes(self):
        return lib().orc_world_nodes(self.h)

    def assemble(self):
        g = np.zeros(3 * self.nodes)
        raw = Bsr()
        lib().orc_world_assemble(self.h, P(g), C.byref(raw))
        return g, OBsr(raw)

    def set_soft_targets(self, t):
        lib().orc_world_set_soft_targets(self.h, P(np.ascontiguousarray(t, np.float64).ravel()))

    @property
    def anim_rate(self):
        return lib().orc_world_get_anim_rate(self.h)

    @anim_rate.setter
    def anim_rate(self, r):
        lib().orc_world_set_anim_rate(self.h, float(r))

    def energy(self):
        """{total, kinetic, elastic, barrier, ground, friction, cloth, affine
        shape, soft targets}"""
        t = np.zeros(9)
        lib().orc_world_energy(self.h, P(t))
        return t

    def newton_iteration(self):
        a = C.c_double()
        it = C.c_int32()
        xd = C.c_double()
        cv = C.c_int32()
        lib().orc_world_newton_iteration(self.h, C.byref(a), C.byref(it), C.byref(xd), C.byref(cv))
        return dict(alpha=a.value, pcg_iterations=it.value, xdelta=xd.value, converged=cv.value)

    def set_state(self, x, v, xt, kappa):
        a = [np.ascontiguousarray(q, np.float64).ravel() for q in (x, v, xt)]
        lib().orc_world_set_state(self.h, P(a[0]), P(a[1]), P(a[2]), float(kappa))

    def advance(self, n):
        st = C.c_int32()
        it = C.c_int32()
        done = lib().orc_world_advance(self.h, int(n), C.byref(st), C.byref(it))
        return dict(newton_iterations=done, time_steps_begun=st.value, pcg_iterations=it.value)

    def ccd_bound(self, dx):
        """step bounds of the vertex direction dx (x_new = x - alpha dx)"""
        out = np.zeros(6)
        lib().orc_world_ccd_bound(self.h, P(np.ascontiguousarray(dx, np.float64).ravel()),
                                  P(out))
        return dict(alpha_ground=out[0], alpha_self=out[1], n_candidates=int(out[2]),
                    max_speed=out[3], alpha_cfl=out[4], alpha=out[5])

    def ccd_candidates(self, dx, alpha, cap=1 << 20):
        out = np.zeros((cap, 4), np.int32)
        n = lib().orc_world_ccd_candidates(
            self.h, P(np.ascontiguousarray(dx, np.float64).ravel()), float(alpha),
            P(out, C.c_int32), cap)
        return out[:min(n, cap)].copy()

    def kappa_info(self):
        out = np.zeros(5)
        lib().orc_world_kappa_info(self.h, P(out))
        return dict(suggest=out[0], upper=out[1], mean_mass=out[2], bbox_diag2=out[3],
                    dhat2=out[4])

    def init_kappa(self):
        return lib().orc_world_init_kappa(self.h)

    def time_step(self):
        n = C.c_int32()
        it = C.c_int32()
        lib().orc_world_time_step(self.h, C.byref(n), C.byref(it))
        return dict(newton_iterations=n.value, pcg_iterations=it.value)


def pt_ccd(p, t0, t1, t2, dp, dt0, dt1, dt2, eta=0.2, thickness=0.0):
    a = [np.ascontiguousarray(v, np.float64) for v in (p, t0, t1, t2, dp, dt0, dt1, dt2)]
    return lib().orc_pt_ccd(*[P(v) for v in a], eta, thickness)


def ee_ccd(a0, a1, b0, b1, da0, da1, db0, db1, eta=0.2, thickness=0.0):
    a = [np.ascontiguousarray(v, np.float64) for v in (a0, a1, b0, b1, da0, da1, db0, db1)]
    return lib().orc_ee_ccd(*[P(v) for v in a], eta, thickness)


def barrier_grad_hess(x, code, dhat2, kappa):
    x = np.ascontiguousarray(x, np.float64)
    code = np.ascontiguousarray(code, np.int32)
    ids = np.zeros(4, np.int32)
    g = np.zeros(12)
    H = np.zeros(144)
    nv = lib().orc_barrier_grad_hess(P(x), P(x), P(code, C.c_int32), dhat2, kappa,
                                     P(ids, C.c_int32), P(g), P(H))
    return ids[:nv], g[:3 * nv], H[:(3 * nv) ** 2].reshape(3 * nv, 3 * nv)


def barrier_energy(x, code, dhat2, kappa):
    x = np.ascontiguousarray(x, np.float64)
    code = np.ascontiguousarray(code, np.int32)
    return lib().orc_barrier_energy(P(x), P(x), P(code, C.c_int32), dhat2, kappa)


def tet_grad_hess(x, tet, dminv, vol, mu, lam, dt):
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(tet, np.uint32)
    Dm = np.ascontiguousarray(dminv, np.float64).reshape(9)
    g = np.zeros(12)
    H = np.zeros(144)
    lib().orc_tet_grad_hess(P(x), P(t, C.c_uint32), P(Dm), vol, mu, lam, dt, P(g), P(H))
    return g, H.reshape(12, 12)


def tet_energy(x, tet, dminv, vol, mu, lam):
    x = np.ascontiguousarray(x, np.float64)
    t = np.ascontiguousarray(tet, np.uint32)
    Dm = np.ascontiguousarray(dminv, np.float64).reshape(9)
    return lib().orc_tet_energy(P(x), P(t, C.c_uint32), P(Dm), vol, mu, lam)


def ground(x, n, off, dhat2, kappa):
    x = np.ascontiguousarray(x, np.float64)
    n = np.ascontiguousarray(n, np.float64)
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_ground_grad_hess(P(x), P(n), off, dhat2, kappa, P(g), P(H))
    d = float(n @ x - off)
    return lib().orc_ground_energy(d * d, dhat2) * kappa, g, H.reshape(3, 3)


# ---------------------------------------------------------------- friction
def _f64(a):
    return np.ascontiguousarray(a, np.float64)


def friction_lag(x, code, dhat2, kappa):
    """(kind, lam, coord[2], basis 3x2) of a contact code at x; kind -1: none."""
    x = _f64(x).ravel()
    c = np.ascontiguousarray(code, np.int32)
    lam = np.zeros(1)
    coord = np.zeros(2)
    basis = np.zeros(6)
    kind = lib().orc_friction_lag(P(x), P(c, C.c_int32), dhat2, kappa, P(lam), P(coord),
                                  P(basis))
    return kind, float(lam[0]), coord, basis.reshape(3, 2)


def friction_energy(x, xp, code, coord, basis, lam, eps2):
    return lib().orc_friction_energy(P(_f64(x).ravel()), P(_f64(xp).ravel()),
                                     P(np.ascontiguousarray(code, np.int32), C.c_int32),
                                     P(_f64(coord)), P(_f64(basis).ravel()), lam, eps2)


def friction_grad_hess(x, xp, code, coord, basis, lam, eps2, coef):
    ids = np.zeros(4, np.int32)
    g = np.zeros(12)
    H = np.zeros(144)
    m = lib().orc_friction_grad_hess(P(_f64(x).ravel()), P(_f64(xp).ravel()),
                                     P(np.ascontiguousarray(code, np.int32), C.c_int32),
                                     P(_f64(coord)), P(_f64(basis).ravel()), lam, eps2, coef,
                                     P(ids, C.c_int32), P(g), P(H))
    n = 3 * m
    return ids[:m].copy(), g[:n].copy(), H[:n * n].reshape(n, n).copy()


def friction_lag_ground(x, n, off, dhat2, kappa):
    return lib().orc_friction_lag_ground(P(_f64(x)), P(_f64(n)), off, dhat2, kappa)


def friction_ground_energy(x, xp, n, lam, eps2):
    return lib().orc_friction_ground_energy(P(_f64(x)), P(_f64(xp)), P(_f64(n)), lam, eps2)


def friction_ground_grad_hess(x, xp, n, lam, eps2, coef):
    g = np.zeros(3)
    H = np.zeros(9)
    lib().orc_friction_ground_grad_hess(P(_f64(x)), P(_f64(xp)), P(_f64(n)), lam, eps2, coef,
                                        P(g), P(H))
    return g, H.reshape(3, 3)


# ------------------------------------------------------------------- cloth
def cloth_params(E, nu, bend_E, thickness, strain_rate):
    """(stretch, shear, bend) stiffnesses"""
    o = [C.c_double() for _ in range(3)]
    lib().orc_cloth_params(E, nu, bend_E, thickness, strain_rate, *[C.byref(v) for v in o])
    return tuple(v.value for v in o)


def tri_precompute(x_rest, tri):
    Dm = np.zeros(4)
    a = np.zeros(1)
    lib().orc_tri_precompute(P(_f64(x_rest).ravel()), P(np.ascontiguousarray(tri, np.uint32),
                             C.c_uint32), P(Dm), P(a))
    return Dm.reshape(2, 2), float(a[0])


def bw_psi(F, stretch, shear, r):
    return lib().orc_bw_psi(P(_f64(F).ravel()), stretch, shear, r)


def bw_pk1(F, stretch, shear, r):
    Pk = np.zeros(6)
    lib().orc_bw_pk1(P(_f64(F).ravel()), stretch, shear, r, P(Pk))
    return Pk.reshape(3, 2)


def bw_hess(F, stretch, shear, r, project=True):
    """6x6 Hessian in vec(F) = (F[:,0], F[:,1]) order"""
    H = np.zeros(36)
    lib().orc_bw_hess(P(_f64(F).ravel()), stretch, shear, r, int(project), P(H))
    return H.reshape(6, 6)


def tri_energy(x, tri, dminv, area, stretch, shear, r):
    return lib().orc_tri_energy(P(_f64(x).ravel()), P(np.ascontiguousarray(tri, np.uint32),
                                C.c_uint32), P(_f64(dminv).ravel()), area, stretch, shear, r)


def tri_grad_hess(x, tri, dminv, area, stretch, shear, r, dt):
    g = np.zeros(9)
    H = np.zeros(81)
    lib().orc_tri_grad_hess(P(_f64(x).ravel()), P(np.ascontiguousarray(tri, np.uint32),
                            C.c_uint32), P(_f64(dminv).ravel()), area, stretch, shear, r, dt,
                            P(g), P(H))
    return g, H.reshape(9, 9)


def dihedral(q0, q1, q2, q3, derivs=True):
    """(theta, grad 12, Hessian 12x12) or theta"""
    q = [_f64(v) for v in (q0, q1, q2, q3)]
    if not derivs:
        return lib().orc_dihedral(*[P(v) for v in q], None, None)
    g = np.zeros(12)
    H = np.zeros(144)
    t = lib().orc_dihedral(*[P(v) for v in q], P(g), P(H))
    return t, g, H.reshape(12, 12)


def bend_energy(x, hinge, theta0, len0, bend):
    return lib().orc_bend_energy(P(_f64(x).ravel()), P(np.ascontiguousarray(hinge, np.uint32),
                                 C.c_uint32), theta0, len0, bend)


def bend_grad_hess(x, hinge, theta0, len0, bend, dt):
    g = np.zeros(12)
    H = np.zeros(144)
    lib().orc_bend_grad_hess(P(_f64(x).ravel()), P(np.ascontiguousarray(hinge, np.uint32),
                             C.c_uint32), theta0, len0, bend, dt, P(g), P(H))
    return g, H.reshape(12, 12)


def bend_hinges(tris):
    t = np.ascontiguousarray(tris, np.uint32)
    n = lib().orc_bend_hinges(int(t.shape[0]), P(t, C.c_uint32), None)
    h = np.zeros((max(n, 1), 4), np.uint32)
    lib().orc_bend_hinges(int(t.shape[0]), P(t, C.c_uint32), P(h, C.c_uint32))
    return h[:n]


# ------------------------------------------------------------- affine bodies
def abd_shape(q):
    """(energy, gradient 9, unprojected Hessian 9x9) of ||A A^T - I||_F^2"""
    q = _f64(q)
    g = np.zeros(9)
    H = np.zeros(81)
    lib().orc_abd_shape_grad(P(q), P(g))
    lib().orc_abd_shape_hess(P(q), P(H))
    return lib().orc_abd_shape_energy(P(q)), g, H.reshape(9, 9)


def abd_jmat(xbar):
    J = np.zeros(36)
    lib().orc_abd_jmat(P(_f64(xbar)), P(J))
    return J.reshape(3, 12)


# ------------------------------------------------------------- warm-start MLP
def mlp_loss_grad(w, X, T, want_y=False):
    fp = C.POINTER(C.c_float)
    w = np.ascontiguousarray(w, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    g = np.zeros(lib().orc_mlp_params())
    Y = np.zeros((X.shape[0], 12), np.float32)
    L = lib().orc_mlp_loss_grad(w.ctypes.data_as(fp), X.ctypes.data_as(fp), T.ctypes.data_as(fp),
                                X.shape[0], P(g), Y.ctypes.data_as(fp))
    return (L, g, Y) if want_y else (L, g)
