"""Run-to-run reproducibility of the FP64 linear solver: the same system
solved twice in one process gives bit-identical MAS applies (fixed-order
bank restriction and Galerkin bank matrices) and PCG iterates (fixed-order
dot products); a Newton iteration of a scene whose contact terms are few
reproduces bit for bit as well.  Contact-rich scenes still sum their contact
gradient and Hessian blocks with atomics, in an order that varies between
runs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(a, b):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_pcg_and_mas_apply_are_bit_reproducible():
    import sgipc
    from sgipc import scene as S
    sc = S.stack_scene(n=3, cells=4, friction_mu=0.2)
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    sim = sgipc.Simulation(sc)
    for _ in range(3):
        sim.newton_iteration()
    g, rp, cols, vals = sim.assemble()
    for prec in (sgipc.SGIPC_MAS_FP64, sgipc.SGIPC_MAS_FP32):
        ls = sgipc.LinearSystem(rp, cols, vals, mas_precision=prec)
        b = -g
        z0, z1 = ls.precond(b), ls.precond(b)
        x0, s0 = ls.pcg(b, rel_tol=1e-8)
        x1, s1 = ls.pcg(b, rel_tol=1e-8)
        print(f"mas_precision={prec} apply equal={_same(z0, z1)} pcg equal={_same(x0, x1)} "
              f"iters {s0.iterations} {s1.iterations}", flush=True)
        assert _same(z0, z1) and _same(x0, x1) and s0.iterations == s1.iterations
        # two systems built from the same matrix give the same bank matrices
        ls2 = sgipc.LinearSystem(rp, cols, vals, mas_precision=prec)
        for bank in (0, ls.mas_info()[2] // 16 - 1):
            m0, i0 = ls.mas_bank(bank)
            m1, i1 = ls2.mas_bank(bank)
            assert _same(m0, m1) and _same(i0, i1)


def test_cloth_newton_iteration_is_bit_reproducible():
    import sgipc
    from sgipc import scene as S
    sc = S.cloth_scene(n=24, n_boxes=2, box_cells=2)
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    sim = sgipc.Simulation(sc)
    sim.set_pcg(rel_tol=1e-6)
    for _ in range(3):
        sim.newton_iteration()
    state = (sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)
    runs = []
    for _ in range(2):
        sim.set_state(*state)
        e = sim.energy()
        g, rp, cols, vals = sim.assemble()
        st = sim.newton_iteration()
        runs.append(dict(energy=e, grad=g, vals=vals, x=sim.positions(),
                         pcg=st["pcg_iterations"]))
    bad = [k for k in runs[0] if not _same(runs[0][k], runs[1][k])]
    print(f"cloth: differing={bad}", flush=True)
    assert not bad
