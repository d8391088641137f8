"""ctypes binding of libsgipc.so (include/sgipc.h).

The library is the only compute path: if it cannot be loaded every call
raises; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .scene import SceneDesc

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libsgipc.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
up = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)

SGIPC_PRECOND_DIAG = 0
SGIPC_PRECOND_MAS = 1
SGIPC_MAS_FP32 = 0
SGIPC_MAS_FP64 = 1
SGIPC_SOLVE_FP64 = 0
SGIPC_SOLVE_MIXED = 1


class PcgCfg(C.Structure):
    _fields_ = [("rel_tol", C.c_double), ("abs_tol", C.c_double), ("max_iter", C.c_int32),
                ("use_preconditioned_norm", C.c_int32), ("check_every", C.c_int32),
                ("solver_precision", C.c_int32)]


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("outer_iterations", C.c_int32),
                ("converged", C.c_int32), ("pad", C.c_int32), ("rel_residual", C.c_double),
                ("rz0", C.c_double), ("rz", C.c_double), ("ms", C.c_double)]


class IterStats(C.Structure):
    _fields_ = [("pcg_iterations", C.c_int32), ("n_pairs", C.c_int32), ("n_ground", C.c_int32),
                ("line_search_steps", C.c_int32), ("converged", C.c_int32),
                ("solver_fallback", C.c_int32), ("warm_start", C.c_int32),
                ("warm_fallback", C.c_int32),
                ("pcg_rel_residual", C.c_double), ("alpha", C.c_double),
                ("alpha_ccd", C.c_double), ("xdelta", C.c_double), ("energy0", C.c_double),
                ("energy1", C.c_double), ("ms_assembly", C.c_double), ("ms_solve", C.c_double),
                ("ms_ccd", C.c_double), ("ms_line_search", C.c_double),
                ("ms_total", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f not in ("pad",)}


class StepStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("pcg_iterations", C.c_int32),
                ("ms_total", C.c_double), ("capped", C.c_int32), ("pad", C.c_int32)]


class AdvanceStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int32), ("time_steps_begun", C.c_int32),
                ("pcg_iterations", C.c_int32), ("solver_fallbacks", C.c_int32),
                ("max_pcg_rel_residual", C.c_double), ("ms_total", C.c_double),
                ("warm_starts", C.c_int32), ("capped_steps", C.c_int32),
                ("ms_assembly", C.c_double), ("ms_solve", C.c_double), ("ms_ccd", C.c_double),
                ("ms_line_search", C.c_double), ("line_search_steps", C.c_int32),
                ("ccd_limited", C.c_int32), ("warm_fallbacks", C.c_int32), ("pad", C.c_int32)]


class SgipcError(RuntimeError):
    pass


_lib = None

# every symbol include/sgipc.h declares (tests check the export table)
EXPORTS = [
    "sgipc_last_error", "sgipc_device_info",
    "sgipc_ls_create", "sgipc_ls_destroy", "sgipc_ls_spmv", "sgipc_ls_precond_apply",
    "sgipc_ls_pcg", "sgipc_ls_mas_info", "sgipc_ls_mas_chain", "sgipc_ls_mas_bank",
    "sgipc_create", "sgipc_destroy", "sgipc_set_positions", "sgipc_get_positions",
    "sgipc_get_velocities", "sgipc_set_xtilde", "sgipc_get_xtilde", "sgipc_get_kappa",
    "sgipc_set_kappa", "sgipc_lbvh_build", "sgipc_contact_query", "sgipc_assemble",
    "sgipc_energy", "sgipc_newton_iteration", "sgipc_time_step", "sgipc_advance",
    "sgipc_profile_enable", "sgipc_profile_query", "sgipc_problem_sizes",
    "sgipc_set_state", "sgipc_set_pcg", "sgipc_learned_mode", "sgipc_warm_start_adaptive",
    "sgipc_learned_samples",
    "sgipc_mlp_train", "sgipc_mlp_loss_grad", "sgipc_mlp_set", "sgipc_mlp_get",
    "sgipc_ccd_bound", "sgipc_ccd_candidates", "sgipc_ls_pcg_noalloc",
    "sgipc_set_soft_targets", "sgipc_get_animation_rate",
    "sgipc_ls_create_bcoo", "sgipc_ls_create_mm", "sgipc_ls_block_rows", "sgipc_mm_read_vector",
    "sgipc_mm_read_upper_bcoo", "sgipc_ls_set_stream", "sgipc_set_dof_order", "sgipc_mm_export",
    "sgipc_set_stream", "sgipc_mm_export_step", "sgipc_step_log", "sgipc_spatial_sort_keys",
]
SGIPC_DOF_FEM_FIRST = 0
SGIPC_DOF_ABD_FIRST = 1
MLP_PARAMS, MLP_IN, MLP_OUT = 21260, 24, 12


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SgipcError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.sgipc_last_error.restype = C.c_char_p
        L.sgipc_device_info.argtypes = [C.c_char_p, C.c_int32]
        L.sgipc_ls_create.argtypes = [C.c_int32, C.c_int32, ip, ip, dp, ip, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_void_p)]
        L.sgipc_ls_destroy.argtypes = [C.c_void_p]
        L.sgipc_ls_spmv.argtypes = [C.c_void_p, dp, dp]
        L.sgipc_ls_precond_apply.argtypes = [C.c_void_p, dp, dp]
        L.sgipc_ls_pcg.argtypes = [C.c_void_p, dp, dp, C.POINTER(PcgCfg), C.POINTER(PcgStats)]
        L.sgipc_ls_pcg_noalloc.argtypes = [C.c_void_p, dp, dp, C.POINTER(PcgCfg),
                                           C.POINTER(PcgStats)]
        L.sgipc_ls_mas_info.argtypes = [C.c_void_p, ip, ip, ip]
        L.sgipc_ls_mas_chain.argtypes = [C.c_void_p, ip]
        L.sgipc_ls_mas_bank.argtypes = [C.c_void_p, C.c_int32, dp, dp]
        L.sgipc_ls_create_bcoo.argtypes = [C.c_int32, C.c_int32, ip, ip, dp, ip, C.c_int32,
                                           C.c_int32, C.POINTER(C.c_void_p)]
        L.sgipc_ls_create_mm.argtypes = [C.c_char_p, ip, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_void_p)]
        L.sgipc_ls_block_rows.argtypes = [C.c_void_p, ip]
        L.sgipc_mm_read_vector.argtypes = [C.c_char_p, dp, C.c_int32, ip]
        L.sgipc_mm_read_upper_bcoo.argtypes = [C.c_char_p, ip, ip, ip, ip, dp, C.c_int32]
        L.sgipc_spatial_sort_keys.argtypes = [C.c_int32, dp, C.c_int32, up, C.c_int32, u64p, up]
        L.sgipc_ls_set_stream.argtypes = [C.c_void_p, C.c_void_p]
        if hasattr(L, "sgipc_create"):
            L.sgipc_create.argtypes = [C.POINTER(SceneDesc), C.POINTER(C.c_void_p)]
            L.sgipc_destroy.argtypes = [C.c_void_p]
            for f in ("sgipc_set_positions", "sgipc_get_positions", "sgipc_get_velocities",
                      "sgipc_set_xtilde", "sgipc_get_xtilde"):
                getattr(L, f).argtypes = [C.c_void_p, dp]
            L.sgipc_get_kappa.argtypes = [C.c_void_p, dp]
            L.sgipc_set_kappa.argtypes = [C.c_void_p, C.c_double]
            L.sgipc_set_state.argtypes = [C.c_void_p, dp, dp, dp, C.c_double]
            L.sgipc_set_pcg.argtypes = [C.c_void_p, C.POINTER(PcgCfg)]
            L.sgipc_lbvh_build.argtypes = [C.c_void_p, C.c_int32, u64p, up, dp]
            L.sgipc_contact_query.argtypes = [C.c_void_p, ip, C.c_int32, ip, up, C.c_int32, ip]
            L.sgipc_assemble.argtypes = [C.c_void_p, dp, ip, ip, ip, dp, C.c_int32]
            L.sgipc_energy.argtypes = [C.c_void_p, dp]
            L.sgipc_set_soft_targets.argtypes = [C.c_void_p, dp]
            L.sgipc_set_dof_order.argtypes = [C.c_void_p, C.c_int32]
            L.sgipc_mm_export.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32]
            L.sgipc_set_stream.argtypes = [C.c_void_p, C.c_void_p]
            L.sgipc_mm_export_step.argtypes = [C.c_void_p, C.c_char_p, C.c_int32]
            L.sgipc_step_log.argtypes = [C.c_void_p, C.POINTER(IterStats), C.c_int32,
                                         C.POINTER(C.c_int32)]
            L.sgipc_get_animation_rate.argtypes = [C.c_void_p, dp]
            L.sgipc_newton_iteration.argtypes = [C.c_void_p, C.POINTER(IterStats)]
            L.sgipc_ccd_bound.argtypes = [C.c_void_p, dp, dp]
            L.sgipc_ccd_candidates.argtypes = [C.c_void_p, dp, C.c_double, ip, C.c_int32, ip]
            L.sgipc_time_step.argtypes = [C.c_void_p, C.POINTER(StepStats)]
            L.sgipc_advance.argtypes = [C.c_void_p, C.c_int32, C.POINTER(AdvanceStats)]
            L.sgipc_profile_enable.argtypes = [C.c_void_p, C.c_int32]
            L.sgipc_profile_query.argtypes = [C.c_void_p, C.c_char_p, ip, dp]
            L.sgipc_problem_sizes.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
            fp = C.POINTER(C.c_float)
            L.sgipc_learned_mode.argtypes = [C.c_void_p, C.c_int32]
            L.sgipc_warm_start_adaptive.argtypes = [C.c_void_p, C.c_int32]
            L.sgipc_learned_samples.argtypes = [C.c_void_p, fp, fp, C.c_int32,
                                                C.POINTER(C.c_int32)]
            L.sgipc_mlp_train.argtypes = [C.c_void_p, fp, fp, C.c_int32, C.c_int32, C.c_double,
                                          C.POINTER(C.c_double)]
            L.sgipc_mlp_loss_grad.argtypes = [C.c_void_p, fp, fp, C.c_int32, fp,
                                              C.POINTER(C.c_double)]
            L.sgipc_mlp_set.argtypes = [C.c_void_p, fp]
            L.sgipc_mlp_get.argtypes = [C.c_void_p, fp]
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise SgipcError(f"sgipc error {rc}: {lib().sgipc_last_error().decode()}")


def P(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def device_info():
    buf = C.create_string_buffer(256)
    check(lib().sgipc_device_info(buf, 256))
    return buf.value.decode()


def read_mm_vector(path):
    """MatrixMarket array real M x 1 (read_mm_array_vector_mx1, main.cu)"""
    m = C.c_int32()
    check(lib().sgipc_mm_read_vector(os.fsencode(path), None, 0, C.byref(m)))
    b = np.zeros(m.value)
    check(lib().sgipc_mm_read_vector(os.fsencode(path), P(b), m.value, C.byref(m)))
    return b


def spatial_sort_keys(x, tets=None):
    """(sorted keys, order) of the vertices x (tets None: Morton codes) or of
    the tets (centroid keys); keys are code << 32 | index (GIPC.cu:279-461)"""
    x = np.ascontiguousarray(x, np.float64).reshape(-1, 3)
    if tets is None:
        kind, n, t = 0, x.shape[0], None
    else:
        t = np.ascontiguousarray(tets, np.uint32).reshape(-1, 4)
        kind, n = 1, t.shape[0]
    keys = np.zeros(n, np.uint64)
    order = np.zeros(n, np.uint32)
    check(lib().sgipc_spatial_sort_keys(kind, P(x), x.shape[0],
                                        None if t is None else P(t, C.c_uint32),
                                        0 if t is None else n, P(keys, C.c_uint64),
                                        P(order, C.c_uint32)))
    return keys, order


def read_mm_upper_bcoo(path):
    """(nb, rows, cols, vals[k,3,3]) of the unique upper blocks a MatrixMarket
    A describes (to_upper_triangle_block_triplets, main.cu)"""
    nb, nnz = C.c_int32(), C.c_int32()
    f = os.fsencode(path)
    check(lib().sgipc_mm_read_upper_bcoo(f, C.byref(nb), C.byref(nnz), None, None, None, 0))
    r = np.zeros(nnz.value, np.int32)
    c = np.zeros(nnz.value, np.int32)
    v = np.zeros((nnz.value, 3, 3))
    check(lib().sgipc_mm_read_upper_bcoo(f, C.byref(nb), C.byref(nnz), P(r, C.c_int32),
                                         P(c, C.c_int32), P(v), nnz.value))
    return nb.value, r, c, v


class LinearSystem:
    """Stand-alone full symmetric BSR system with MAS or block-Jacobi PCG
    (counterpart of GlobalLinearSystem::solve_loaded_system)."""

    def __init__(self, rowptr, cols, vals, part_id=None, precond=SGIPC_PRECOND_MAS,
                 mas_precision=SGIPC_MAS_FP32, _handle=None, _nb=None):
        self.part = None if part_id is None else np.ascontiguousarray(part_id, np.int32)
        if _handle is not None:
            self.h, self.nb = _handle, _nb
            return
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.cols = np.ascontiguousarray(cols, np.int32)
        self.vals = np.ascontiguousarray(vals, np.float64).reshape(-1)
        self.nb = self.rowptr.size - 1
        h = C.c_void_p()
        check(lib().sgipc_ls_create(self.nb, self.cols.size, P(self.rowptr, C.c_int32),
                                    P(self.cols, C.c_int32), P(self.vals),
                                    None if self.part is None else P(self.part, C.c_int32),
                                    precond, mas_precision, C.byref(h)))
        self.h = h

    @classmethod
    def _wrap(cls, h, part_id):
        nb = C.c_int32()
        check(lib().sgipc_ls_block_rows(h, C.byref(nb)))
        return cls(None, None, None, part_id, _handle=h, _nb=nb.value)

    @classmethod
    def from_bcoo(cls, nb, rows, cols, vals, part_id=None, precond=SGIPC_PRECOND_MAS,
                  mas_precision=SGIPC_MAS_FP32):
        """from the reference's merged upper BCOO (row <= col, row-major 3x3)"""
        r = np.ascontiguousarray(rows, np.int32)
        c = np.ascontiguousarray(cols, np.int32)
        v = np.ascontiguousarray(vals, np.float64).reshape(-1)
        pid = None if part_id is None else np.ascontiguousarray(part_id, np.int32)
        h = C.c_void_p()
        check(lib().sgipc_ls_create_bcoo(int(nb), r.size, P(r, C.c_int32), P(c, C.c_int32),
                                         P(v), None if pid is None else P(pid, C.c_int32),
                                         precond, mas_precision, C.byref(h)))
        return cls._wrap(h, pid)

    @classmethod
    def from_matrix_market(cls, A_path, part_id=None, precond=SGIPC_PRECOND_MAS,
                           mas_precision=SGIPC_MAS_FP32):
        """solve-only mode: A from a MatrixMarket file (main.cu --A)"""
        pid = None if part_id is None else np.ascontiguousarray(part_id, np.int32)
        h = C.c_void_p()
        check(lib().sgipc_ls_create_mm(os.fsencode(A_path),
                                       None if pid is None else P(pid, C.c_int32), precond,
                                       mas_precision, C.byref(h)))
        return cls._wrap(h, pid)

    def close(self):
        if getattr(self, "h", None):
            lib().sgipc_ls_destroy(self.h)
            self.h = None

    __del__ = close

    def spmv(self, x):
        y = np.zeros(3 * self.nb)
        check(lib().sgipc_ls_spmv(self.h, P(np.ascontiguousarray(x, np.float64)), P(y)))
        return y

    def precond(self, r):
        z = np.zeros(3 * self.nb)
        check(lib().sgipc_ls_precond_apply(self.h, P(np.ascontiguousarray(r, np.float64)), P(z)))
        return z

    def pcg(self, b, rel_tol=1e-4, abs_tol=0.0, max_iter=0, use_rz=True, check_every=10,
            solver_precision=SGIPC_SOLVE_FP64):
        x = np.zeros(3 * self.nb)
        cfg = PcgCfg(rel_tol, abs_tol, max_iter, int(use_rz), check_every, solver_precision)
        st = PcgStats()
        check(lib().sgipc_ls_pcg(self.h, P(np.ascontiguousarray(b, np.float64)), P(x),
                                 C.byref(cfg), C.byref(st)))
        return x, st

    def pcg_noalloc(self, b, rel_tol=1e-4, abs_tol=0.0, max_iter=0, use_rz=True,
                    check_every=10):
        """the FP64 solve of pcg() with no device allocation (static workspace)"""
        x = np.zeros(3 * self.nb)
        cfg = PcgCfg(rel_tol, abs_tol, max_iter, int(use_rz), check_every, SGIPC_SOLVE_FP64)
        st = PcgStats()
        check(lib().sgipc_ls_pcg_noalloc(self.h, P(np.ascontiguousarray(b, np.float64)), P(x),
                                         C.byref(cfg), C.byref(st)))
        return x, st

    def mas_info(self):
        L = C.c_int32()
        tc = C.c_int32()
        sizes = np.zeros(16, np.int32)
        check(lib().sgipc_ls_mas_info(self.h, C.byref(L), P(sizes, C.c_int32), C.byref(tc)))
        return L.value, sizes[: 2 * (L.value + 1)].reshape(-1, 2), tc.value

    def mas_chain(self):
        L = self.mas_info()[0]
        c = np.zeros(self.nb * (L - 1), np.int32)
        check(lib().sgipc_ls_mas_chain(self.h, P(c, C.c_int32)))
        return c.reshape(self.nb, L - 1)

    def mas_bank(self, bank):
        """(48x48 assembled bank matrix, the inverse the apply uses), FP64"""
        m = np.zeros((48, 48))
        inv = np.zeros((48, 48))
        check(lib().sgipc_ls_mas_bank(self.h, int(bank), P(m), P(inv)))
        return m, inv


class Simulation:
    """Device-resident Stiff-GIPC state for one scene (GIPC counterpart)."""

    def __init__(self, scene):
        self.scene = scene
        self._desc = scene.desc()
        h = C.c_void_p()
        check(lib().sgipc_create(C.byref(self._desc), C.byref(h)))
        self.h = h
        self.nv = scene.n_verts

    def close(self):
        if getattr(self, "h", None):
            lib().sgipc_destroy(self.h)
            self.h = None

    __del__ = close

    def _get(self, f):
        a = np.zeros(3 * self.nv)
        check(f(self.h, P(a)))
        return a.reshape(-1, 3)

    def positions(self):
        return self._get(lib().sgipc_get_positions)

    def velocities(self):
        return self._get(lib().sgipc_get_velocities)

    def xtilde(self):
        return self._get(lib().sgipc_get_xtilde)

    def set_positions(self, x):
        check(lib().sgipc_set_positions(self.h, P(np.ascontiguousarray(x, np.float64).ravel())))

    def set_xtilde(self, xt):
        check(lib().sgipc_set_xtilde(self.h, P(np.ascontiguousarray(xt, np.float64).ravel())))

    @property
    def kappa(self):
        k = C.c_double()
        check(lib().sgipc_get_kappa(self.h, C.byref(k)))
        return k.value

    @kappa.setter
    def kappa(self, k):
        check(lib().sgipc_set_kappa(self.h, float(k)))

    def lbvh(self, which):
        n = self.scene.faces.shape[0] if which == 0 else self.scene.edges.shape[0]
        keys = np.zeros(n, np.uint64)
        nodes = np.zeros((2 * n - 1, 4), np.uint32)
        aabbs = np.zeros((2 * n - 1, 6), np.float64)
        check(lib().sgipc_lbvh_build(self.h, which, P(keys, C.c_uint64), P(nodes, C.c_uint32),
                                     P(aabbs)))
        return keys, nodes, aabbs

    def contacts(self, cap=None):
        cap = cap or max(1000, 4 * (self.scene.surf_verts.size + self.scene.edges.shape[0]))
        pairs = np.zeros((cap, 4), np.int32)
        counts = np.zeros(5, np.int32)
        gcap = max(1, self.scene.surf_verts.size)
        ground = np.zeros(gcap, np.uint32)
        ng = C.c_int32()
        check(lib().sgipc_contact_query(self.h, P(pairs, C.c_int32), cap, P(counts, C.c_int32),
                                        P(ground, C.c_uint32), gcap, C.byref(ng)))
        return pairs[: counts[0]].copy(), counts, ground[: ng.value].copy()

    @property
    def nodes(self):
        """system block rows: FEM vertices, then 4 per affine body"""
        return int(self.sizes()["n_nodes"])

    def assemble(self):
        nn = self.nodes
        g = np.zeros(3 * nn)
        nnz = C.c_int32()
        check(lib().sgipc_assemble(self.h, P(g), C.byref(nnz), None, None, None, 0))
        rowptr = np.zeros(nn + 1, np.int32)
        cols = np.zeros(nnz.value, np.int32)
        vals = np.zeros((nnz.value, 3, 3))
        check(lib().sgipc_assemble(self.h, P(g), C.byref(nnz), P(rowptr, C.c_int32),
                                   P(cols, C.c_int32), P(vals), nnz.value))
        return g, rowptr, cols, vals

    def energy(self):
        """{total, kinetic, elastic, barrier, ground, friction, cloth, affine
        shape, soft targets}"""
        t = np.zeros(9)
        check(lib().sgipc_energy(self.h, P(t)))
        return t

    def set_dof_order(self, order):
        """SGIPC_DOF_FEM_FIRST (default) or SGIPC_DOF_ABD_FIRST (the reference's
        GlobalLinearSystem layout) for assemble() and export_matrix_market()"""
        check(lib().sgipc_set_dof_order(self.h, int(order)))

    def export_matrix_market(self, out_dir, frame=0, newton_step=0):
        """A and b of the Newton system at the current positions, written as
        out_dir/linear_system/frame_<frame:05d>_ns_<newton_step>_{A,b}.mtx"""
        check(lib().sgipc_mm_export(self.h, os.fsencode(str(out_dir)), int(frame),
                                    int(newton_step)))
        base = os.path.join(str(out_dir), "linear_system", f"frame_{frame:05d}_ns_{newton_step}")
        return base + "_A.mtx", base + "_b.mtx"

    def export_next_step(self, out_dir, frame):
        """export the system of every solving Newton iteration of the next
        time step (frame_<frame>_ns_<k>_{A,b}.mtx under out_dir/linear_system)"""
        check(lib().sgipc_mm_export_step(self.h, os.fsencode(str(out_dir)), int(frame)))

    def step_log(self):
        """per-Newton-iteration statistics of the last time step"""
        n = C.c_int32()
        check(lib().sgipc_step_log(self.h, None, 0, C.byref(n)))
        buf = (IterStats * max(n.value, 1))()
        check(lib().sgipc_step_log(self.h, buf, n.value, C.byref(n)))
        return [buf[i].as_dict() for i in range(n.value)]

    def set_soft_targets(self, t):
        check(lib().sgipc_set_soft_targets(self.h, P(np.ascontiguousarray(t, np.float64).ravel())))

    @property
    def animation_rate(self):
        r = C.c_double()
        check(lib().sgipc_get_animation_rate(self.h, C.byref(r)))
        return r.value

    def newton_iteration(self):
        st = IterStats()
        check(lib().sgipc_newton_iteration(self.h, C.byref(st)))
        return st.as_dict()

    def ccd_bound(self, dx):
        """step bounds of the vertex direction dx (x_new = x - alpha dx)"""
        out = np.zeros(6)
        d = np.ascontiguousarray(dx, np.float64).ravel()
        check(lib().sgipc_ccd_bound(self.h, d.ctypes.data_as(dp), out.ctypes.data_as(dp)))
        return dict(alpha_ground=out[0], alpha_self=out[1], n_candidates=int(out[2]),
                    max_speed=out[3], alpha_cfl=out[4], alpha=out[5])

    def ccd_candidates(self, dx, alpha, cap=1 << 20):
        out = np.zeros((cap, 4), np.int32)
        n = C.c_int32()
        d = np.ascontiguousarray(dx, np.float64).ravel()
        check(lib().sgipc_ccd_candidates(self.h, d.ctypes.data_as(dp), float(alpha),
                                         out.ctypes.data_as(ip), cap, C.byref(n)))
        return out[:min(n.value, cap)].copy()

    def time_step(self):
        st = StepStats()
        check(lib().sgipc_time_step(self.h, C.byref(st)))
        return dict(newton_iterations=st.newton_iterations, pcg_iterations=st.pcg_iterations,
                    ms_total=st.ms_total, capped=st.capped)

    def advance(self, n):
        st = AdvanceStats()
        check(lib().sgipc_advance(self.h, int(n), C.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_ if f != "pad"}

    def set_state(self, x, v, xt, kappa):
        """State at a time-step boundary (x_prev = x), like the oracle's."""
        a = [np.ascontiguousarray(q, np.float64).ravel() for q in (x, v, xt)]
        check(lib().sgipc_set_state(self.h, P(a[0]), P(a[1]), P(a[2]), float(kappa)))

    def set_pcg(self, rel_tol=1e-4, abs_tol=0.0, max_iter=0, use_rz=True, check_every=10,
                solver_precision=SGIPC_SOLVE_FP64):
        cfg = PcgCfg(rel_tol, abs_tol, max_iter, int(use_rz), check_every, solver_precision)
        check(lib().sgipc_set_pcg(self.h, C.byref(cfg)))

    def profile(self, every=1):
        """Time every `every`-th launch of each profiled kernel (0: off)."""
        check(lib().sgipc_profile_enable(self.h, int(every)))

    def profile_query(self, name):
        n = C.c_int32()
        ms = C.c_double()
        check(lib().sgipc_profile_query(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # ---- learned coarse warm start
    def learned_mode(self, mode):
        check(lib().sgipc_learned_mode(self.h, int(mode)))

    def warm_start_adaptive(self, on):
        """Pause warm starts that cost more than they save (default on)."""
        check(lib().sgipc_warm_start_adaptive(self.h, int(bool(on))))

    def learned_samples(self, cap=1 << 16):
        f = np.zeros((cap, MLP_IN), np.float32)
        t = np.zeros((cap, MLP_OUT), np.float32)
        n = C.c_int32()
        check(lib().sgipc_learned_samples(self.h, P(f, C.c_float), P(t, C.c_float), cap,
                                          C.byref(n)))
        return f[:n.value].copy(), t[:n.value].copy()

    def mlp_train(self, feat, target, epochs=200, lr=1e-3):
        f = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        loss = C.c_double()
        check(lib().sgipc_mlp_train(self.h, P(f, C.c_float), P(t, C.c_float), f.shape[0],
                                    int(epochs), float(lr), C.byref(loss)))
        return loss.value

    def mlp_loss_grad(self, feat, target):
        f = np.ascontiguousarray(feat, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        g = np.zeros(MLP_PARAMS, np.float32)
        loss = C.c_double()
        check(lib().sgipc_mlp_loss_grad(self.h, P(f, C.c_float), P(t, C.c_float), f.shape[0],
                                        P(g, C.c_float), C.byref(loss)))
        return loss.value, g

    def mlp_weights(self, w=None):
        if w is not None:
            w = np.ascontiguousarray(w, np.float32)
            assert w.size == MLP_PARAMS
            check(lib().sgipc_mlp_set(self.h, P(w, C.c_float)))
            return w
        out = np.zeros(MLP_PARAMS, np.float32)
        check(lib().sgipc_mlp_get(self.h, P(out, C.c_float)))
        return out

    def sizes(self):
        out = (C.c_int64 * 8)()
        check(lib().sgipc_problem_sizes(self.h, out))
        keys = ["n_nodes", "n_tets", "nnzb", "coo_blocks", "mas_banks", "mas_levels", "pairs",
                "ground"]
        return dict(zip(keys, list(out)))
