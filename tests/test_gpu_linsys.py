"""GPU parity of the linear-system path (BSR SpMV, MAS, PCG) against the
oracle on assembled Stiff-GIPC systems."""
import numpy as np
import pytest

from helpers import assembled_system, rel

pytestmark = pytest.mark.gpu


def _scenes():
    from sgipc import scene as S
    return {
        "bar": S.bar_scene(),
        "cubes8": S.cubes_scene(n_side=2, cells=4, size=0.2, spacing=0.35, seed=3),
    }


@pytest.fixture(scope="module")
def systems(oracle):
    out = {}
    for name, sc in _scenes().items():
        w, g, A = assembled_system(sc, perturb=0.004, seed=1)
        out[name] = (sc, g, A)
    return out


@pytest.mark.parametrize("name", ["bar", "cubes8"])
def test_spmv_matches_oracle(systems, name):
    import sgipc
    sc, g, A = systems[name]
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, sgipc.SGIPC_PRECOND_DIAG)
    rng = np.random.default_rng(7)
    x = rng.normal(size=3 * A.nb)
    y = ls.spmv(x)
    print(f"[spmv] {name} nb={A.nb} nnzb={A.nnzb}", flush=True)
    assert rel(y, A.spmv(x)) < 1e-13


@pytest.mark.parametrize("name", ["bar", "cubes8"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mas_hierarchy_and_apply(systems, oracle, name, prec):
    import sgipc
    sc, g, A = systems[name]
    fp32 = prec == "fp32"
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP32 if fp32 else sgipc.SGIPC_MAS_FP64)
    m = oracle.Mas(A, sc.part_id, fp32=fp32)
    L, sizes, tc = ls.mas_info()
    print(f"[mas] {name} {prec} levels={L} sizes={sizes.tolist()} total={tc}", flush=True)
    assert L == m.levels
    assert tc == m.total_clusters()
    np.testing.assert_array_equal(sizes, m.level_sizes())
    np.testing.assert_array_equal(ls.mas_chain(), m.chain())
    rng = np.random.default_rng(11)
    r = rng.normal(size=3 * A.nb)
    z = ls.precond(r)
    zo = m.apply(r)
    assert rel(z, zo) < (2e-5 if fp32 else 1e-10)


@pytest.mark.parametrize("name", ["bar", "cubes8"])
@pytest.mark.parametrize("pc", ["mas", "diag"])
def test_pcg_matches_oracle(systems, oracle, name, pc):
    import sgipc
    sc, g, A = systems[name]
    pcid = sgipc.SGIPC_PRECOND_MAS if pc == "mas" else sgipc.SGIPC_PRECOND_DIAG
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, pcid, sgipc.SGIPC_MAS_FP64)
    x, st = ls.pcg(g, rel_tol=1e-4)
    m = oracle.Mas(A, sc.part_id, fp32=False) if pc == "mas" else None
    xo, ito, rz0, rz = oracle.pcg(A, g, m, 1 if pc == "mas" else 0, rel_tol=1e-4)
    print(f"[pcg] {name} {pc} it={st.iterations} oracle_it={ito} relres={st.rel_residual:.3e}",
          flush=True)
    assert st.iterations == ito
    assert rel(x, xo) < 1e-8
    D = A.dense()
    assert np.linalg.norm(D @ x - g) / np.linalg.norm(g) == pytest.approx(st.rel_residual, rel=1e-6)


@pytest.fixture(scope="module")
def deep(oracle):
    from helpers import deep_hierarchy_system
    return deep_hierarchy_system()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mas_deep_hierarchy(deep, oracle, prec):
    """Two large coarse levels (own launches) then a small tail restricted in
    one workgroup: hierarchy equal, apply equal to the oracle."""
    import sgipc
    A, part, b = deep
    fp32 = prec == "fp32"
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, part, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP32 if fp32 else sgipc.SGIPC_MAS_FP64)
    m = oracle.Mas(A, part, fp32=fp32)
    L, sizes, tc = ls.mas_info()
    print(f"[mas-deep] {prec} levels={L} sizes={sizes.tolist()}", flush=True)
    np.testing.assert_array_equal(sizes, m.level_sizes())
    banks = [(int(c) + 15) // 16 for c, _ in sizes[1:L]]
    assert banks[0] > 64 and banks[1] > 64 and min(banks[2:]) <= 64
    r = np.random.default_rng(5).normal(size=3 * A.nb)
    assert rel(ls.precond(r), m.apply(r)) < (2e-5 if fp32 else 1e-10)


def test_pcg_deep_hierarchy(deep, oracle):
    import sgipc
    A, part, b = deep
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, part, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP64)
    x, st = ls.pcg(b, rel_tol=1e-6)
    m = oracle.Mas(A, part, fp32=False)
    xo, ito, _, _ = oracle.pcg(A, b, m, 1, rel_tol=1e-6)
    print(f"[pcg-deep] it={st.iterations} oracle_it={ito} relres={st.rel_residual:.3e}",
          flush=True)
    assert st.iterations == ito
    assert rel(x, xo) < 1e-8


@pytest.mark.parametrize("name", ["bar", "cubes8"])
@pytest.mark.parametrize("pc", ["mas", "diag"])
def test_mixed_precision_refinement(systems, oracle, name, pc):
    """FP32 matrix/vectors/MAS PCG with its periodic residual recomputation in
    FP64 (reliable updates) reaches the FP64 true relative residual asked for,
    in about the FP64 solve's iteration count, and its solution agrees with the
    dense FP64 solve (condition number permitting)."""
    import sgipc
    sc, g, A = systems[name]
    pcid = sgipc.SGIPC_PRECOND_MAS if pc == "mas" else sgipc.SGIPC_PRECOND_DIAG
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, pcid, sgipc.SGIPC_MAS_FP32)
    x, st = ls.pcg(g, rel_tol=1e-6, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_MIXED)
    D = A.dense()
    res = np.linalg.norm(D @ x - g) / np.linalg.norm(g)
    x64, st64 = ls.pcg(g, rel_tol=1e-6, use_rz=False)
    print(f"[mixed] {name} {pc} inner_it={st.iterations} outer={st.outer_iterations} "
          f"relres={st.rel_residual:.3e} (dense {res:.3e}); fp64 it={st64.iterations} "
          f"relres={st64.rel_residual:.3e}", flush=True)
    assert st.converged
    assert st.rel_residual <= 1e-6
    assert st.iterations <= 2 * st64.iterations + 20
    assert res == pytest.approx(st.rel_residual, rel=1e-3, abs=1e-12)
    xd = np.linalg.solve(D, g)
    assert rel(x, xd) < 1e-3


def test_mixed_precision_deep_hierarchy(deep, oracle):
    import sgipc
    A, part, b = deep
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, part, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP32)
    x, st = ls.pcg(b, rel_tol=1e-8, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_MIXED)
    print(f"[mixed-deep] it={st.iterations} outer={st.outer_iterations} "
          f"relres={st.rel_residual:.3e}", flush=True)
    assert st.converged and st.rel_residual <= 1e-8
    assert rel(A.spmv(x), b) <= 1.01e-8


def test_mixed_precision_stiff_contrast(systems, oracle):
    """A system with a 1e8 stiffness contrast (one cube's rows and columns
    scaled by 1e4, i.e. D A D): the FP32 iteration may stall or diverge; the
    mixed solve must still reach the FP64 true residual asked for (FP64
    continuation from its iterate), within the FP32 plus FP64 iteration
    budget, and agree with the dense solve"""
    import sgipc
    sc, g, A = systems["cubes8"]
    nb = A.nb
    scale = np.where(sc.body_id == 0, 1e4, 1.0)
    rows = np.repeat(np.arange(nb), np.diff(A.rowptr))
    vals = A.vals.reshape(-1, 3, 3) * (scale[rows] * scale[A.cols])[:, None, None]
    ls = sgipc.LinearSystem(A.rowptr, A.cols, vals.reshape(A.vals.shape), sc.part_id,
                            sgipc.SGIPC_PRECOND_MAS, sgipc.SGIPC_MAS_FP32)
    b = g * np.repeat(scale, 3)
    x, st = ls.pcg(b, rel_tol=1e-10, use_rz=False, solver_precision=sgipc.SGIPC_SOLVE_MIXED)
    x64, st64 = ls.pcg(b, rel_tol=1e-10, use_rz=False)
    print(f"[mixed-stiff] it={st.iterations} outer={st.outer_iterations} "
          f"relres={st.rel_residual:.3e}; fp64 it={st64.iterations} "
          f"relres={st64.rel_residual:.3e}", flush=True)
    assert st.rel_residual <= 1e-10
    assert st64.rel_residual <= 1e-10
    assert st.iterations <= 4 * st64.iterations + 400
    assert rel(x, x64) < 1e-5


@pytest.mark.parametrize("name", ["bar", "cubes8"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_mas_bank_inverse_bit_exact(systems, oracle, name, prec):
    """Every bank inverse of every level equals the oracle's unpivoted
    Gauss-Jordan (the reference's operation order) of the GPU's own bank
    matrix bit for bit (FP64), or its FP32 rounding; the bank matrices equal
    the oracle's Galerkin banks to 1e-12."""
    import sgipc
    sc, g, A = systems[name]
    fp32 = prec == "fp32"
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP32 if fp32 else sgipc.SGIPC_MAS_FP64)
    L, sizes, tc = ls.mas_info()
    nbank = tc // 16
    print(f"[bank inverse] {name} {prec} banks={nbank}", flush=True)
    checked = 0
    for b in range(nbank):
        m, inv = ls.mas_bank(b)
        M = m.copy()
        d = np.arange(48)
        zero = M[d, d] == 0.0
        M[d[zero], d[zero]] = 1.0
        Mi = oracle.inverse_gj(M)
        up = np.triu(Mi)
        sym = up + np.triu(Mi, 1).T
        expect = sym.astype(np.float32).astype(np.float64) if fp32 else sym
        assert np.array_equal(inv.view(np.uint64), expect.view(np.uint64)), f"bank {b}"
        checked += 1
    assert checked == nbank > 0
