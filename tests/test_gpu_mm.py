"""MatrixMarket export of the Newton system (export_matrix_market_files,
global_linear_system.cu:98-261) and the solve-only mode (main.cu:387-646,
1047-1087) on the GPU: the exported A and b, read back by scipy.io, equal the
assembled system bit for bit (upper blocks, their transposes below); the
reference's [ABD | FEM] DOF order is the same system permuted; a system
created from the file, or from its upper BCOO, solves as the BSR-created one
does and as the oracle's PCG on the same matrix does."""
import numpy as np
import pytest
import scipy.io

from helpers import rel

pytestmark = pytest.mark.gpu


def _dense(nn, rowptr, cols, vals):
    A = np.zeros((3 * nn, 3 * nn))
    for i in range(nn):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[3 * i:3 * i + 3, 3 * cols[k]:3 * cols[k] + 3] += vals[k]
    return A


def _upper_block_mask(n):
    bi = np.arange(3 * n) // 3
    return bi[:, None] <= bi[None, :]


def test_export_and_solve_only(tmp_path, oracle):
    import sgipc
    from sgipc import scene as S
    sc = S.abd_scene(n_abd=2, gap=4e-4)
    sc.kappa = 1e4
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    sim = sgipc.Simulation(sc)
    nn, nfem = sim.nodes, sc.n_verts - int((sc.abd_body >= 0).sum())
    A_path, b_path = sim.export_matrix_market(tmp_path, frame=3, newton_step=1)
    assert A_path.endswith("frame_00003_ns_1_A.mtx")
    g, rp, cl, vl = sim.assemble()
    D = _dense(nn, rp, cl, vl)
    A = scipy.io.mmread(A_path).toarray()
    b = scipy.io.mmread(b_path)[:, 0]
    up = _upper_block_mask(nn)
    print(f"[mm] nodes={nn} nnz={scipy.io.mminfo(A_path)[2]}", flush=True)
    assert A.shape == (3 * nn, 3 * nn)
    # the contact blocks are summed by atomics: two assemblies agree to rounding
    assert rel(A[up], D[up]) < 1e-13 and rel(b, g) < 1e-13
    # every block below the diagonal is the exact transpose of its upper block
    assert np.array_equal(np.where(up, 0, A), np.where(up, A, 0).T * (~up))
    # [ABD | FEM]: the same system with the body nodes first
    sim.set_dof_order(sgipc.SGIPC_DOF_ABD_FIRST)
    A2_path, b2_path = sim.export_matrix_market(tmp_path, frame=3, newton_step=2)
    A2 = scipy.io.mmread(A2_path).toarray()
    b2 = scipy.io.mmread(b2_path)[:, 0]
    perm = np.r_[np.arange(nfem, nn), np.arange(nfem)]          # new node -> old node
    pd = (3 * perm[:, None] + np.arange(3)).ravel()
    assert rel(A2, A[np.ix_(pd, pd)]) < 1e-13 and rel(b2, b[pd]) < 1e-13
    # solve-only mode on the exported file (FEM-first) against the BSR route,
    # the upper-BCOO route and the oracle's PCG on the same matrix
    bm = sgipc.read_mm_vector(b_path)
    assert np.array_equal(bm, b)
    ls_mm = sgipc.LinearSystem.from_matrix_market(A_path, None, sgipc.SGIPC_PRECOND_MAS,
                                                  sgipc.SGIPC_MAS_FP64)
    assert ls_mm.nb == nn
    x1, s1 = ls_mm.pcg(bm, rel_tol=1e-6)
    nb, r, c, v = sgipc.read_mm_upper_bcoo(A_path)
    ls_b = sgipc.LinearSystem.from_bcoo(nb, r, c, v, None, sgipc.SGIPC_PRECOND_MAS,
                                        sgipc.SGIPC_MAS_FP64)
    x2, s2 = ls_b.pcg(bm, rel_tol=1e-6)
    Ao = oracle.bsr_from_upper_triplets(nb, r, c, v.reshape(-1, 9))
    m = oracle.Mas(Ao, np.arange(nn, dtype=np.int32) // 16, fp32=False)
    xo, ito, _, _ = oracle.pcg(Ao, bm, m, 1, rel_tol=1e-6)
    print(f"[mm] solve-only iters mm={s1.iterations} bcoo={s2.iterations} oracle={ito} "
          f"relres={s1.rel_residual:.2e}", flush=True)
    assert s1.iterations == s2.iterations == ito
    assert rel(x1, xo) < 1e-8 and rel(x2, xo) < 1e-8
    assert np.linalg.norm(A @ x1 - bm) / np.linalg.norm(bm) == pytest.approx(s1.rel_residual,
                                                                             rel=1e-6)
