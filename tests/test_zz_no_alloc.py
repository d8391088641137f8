"""sgipc_ls_pcg_noalloc (the allocation-free variant of sgipc_ls_pcg, whose
work memory is a fixed device array) against sgipc_ls_pcg and the oracle:
same iteration counts, solutions within 1e-8 of each other and of the
oracle's PCG (the FP64 bar; two runs of either entry point differ at the
1e-10 level because the PCG dot products are summed by FP64 atomics in
arrival order), on two calls in a row with the second system smaller;
sizes beyond the maximum and the mixed-precision mode are refused."""
import ctypes as C

import numpy as np
import pytest

from helpers import assembled_system, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def systems(oracle):
    from sgipc import scene as S
    out = {}
    for name, sc in (("cubes8", S.cubes_scene(n_side=2, cells=4, size=0.2, spacing=0.35, seed=3)),
                     ("bar", S.bar_scene())):
        w, g, A = assembled_system(sc, perturb=0.004, seed=1)
        out[name] = (sc, g, A)
    return out


@pytest.mark.parametrize("pc", ["mas", "diag"])
def test_noalloc_matches_pcg_two_calls(systems, oracle, pc):
    import sgipc
    pcid = sgipc.SGIPC_PRECOND_MAS if pc == "mas" else sgipc.SGIPC_PRECOND_DIAG
    sizes = []
    for name in ("cubes8", "bar"):  # the second system is the smaller one
        sc, g, A = systems[name]
        ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, pcid,
                                sgipc.SGIPC_MAS_FP64)
        x0, s0 = ls.pcg(g, rel_tol=1e-4)
        x1, s1 = ls.pcg_noalloc(g, rel_tol=1e-4)
        m = oracle.Mas(A, sc.part_id, fp32=False) if pc == "mas" else None
        xo, ito, _, _ = oracle.pcg(A, g, m, 1 if pc == "mas" else 0, rel_tol=1e-4)
        print(f"[noalloc] {name} {pc} nb={A.nb} it={s0.iterations}/{s1.iterations} "
              f"oracle={ito} diff={rel(x1, x0):.2e}", flush=True)
        assert s1.iterations == s0.iterations == ito
        assert s1.converged == s0.converged == 1
        assert rel(x1, x0) < 1e-8
        assert rel(x1, xo) < 1e-8
        assert s1.rel_residual == pytest.approx(s0.rel_residual, rel=1e-6)
        sizes.append(A.nb)
    assert sizes[1] < sizes[0]


def test_noalloc_refuses_large_and_mixed(systems):
    import sgipc
    from sgipc._lib import lib, PcgCfg, PcgStats, P
    nb = 65537  # one block row beyond SGIPC_NOALLOC_MAX_ROWS
    rowptr = np.arange(nb + 1, dtype=np.int32)
    cols = np.arange(nb, dtype=np.int32)
    vals = np.tile(np.eye(3).ravel(), nb)
    big = sgipc.LinearSystem(rowptr, cols, vals, None, sgipc.SGIPC_PRECOND_DIAG)
    with pytest.raises(sgipc.SgipcError, match="maximum"):
        big.pcg_noalloc(np.ones(3 * nb))
    sc, g, A = systems["bar"]
    ls = sgipc.LinearSystem(A.rowptr, A.cols, A.vals, sc.part_id, sgipc.SGIPC_PRECOND_MAS,
                            sgipc.SGIPC_MAS_FP64)
    x = np.zeros(3 * A.nb)
    cfg = PcgCfg(1e-4, 0.0, 0, 1, 10, sgipc.SGIPC_SOLVE_MIXED)
    st = PcgStats()
    rc = lib().sgipc_ls_pcg_noalloc(ls.h, P(np.ascontiguousarray(g)), P(x), C.byref(cfg),
                                    C.byref(st))
    assert rc == -1 and b"FP64" in lib().sgipc_last_error()
    # the refusals leave the entry point working
    x1, s1 = ls.pcg_noalloc(g)
    x0, s0 = ls.pcg(g)
    assert s1.iterations == s0.iterations and rel(x1, x0) < 1e-8
