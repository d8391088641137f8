"""GPU parity of the affine-body system (abd_system/) against the oracle.

The HIP path keeps the kinetic term at vertex level and maps contact blocks
to body nodes in the block sink; the oracle forms M (q - q~) and the
J^T H J node blocks from vertex triplets the way the reference's triplet
conversion does.  Both routes must give the same node system, energies,
Newton iterates and trajectories."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _scene(**kw):
    import sgipc
    from sgipc import scene as S
    sc = S.abd_scene(**kw)
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    return sc


def _dense(nn, rowptr, cols, vals):
    A = np.zeros((3 * nn, 3 * nn))
    for i in range(nn):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[3 * i:3 * i + 3, 3 * cols[k]:3 * cols[k] + 3] += vals[k]
    return A


def test_abd_assembly_and_energy(oracle):
    """affine boxes in contact with each other and with a soft box, bodies
    rotated and sheared off their rest shape: node gradient and node Hessian
    (body blocks, J^T H J contact blocks, ABD-FEM coupling) within 1e-10"""
    import sgipc
    sc = _scene(n_abd=2, gap=4e-4)
    sc.kappa = 1e4
    x = sc.x.copy()
    abd = sc.abd_body
    rng = np.random.default_rng(3)
    for b in (0, 1):
        c = x[abd == b].mean(0)
        A = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
        x[abd == b] = c + (x[abd == b] - c) @ A.T + 1e-4 * rng.normal(size=3)
    fem = abd < 0
    x[fem] += 5e-5 * rng.normal(size=(fem.sum(), 3))
    sim = sgipc.Simulation(sc)
    sim.set_positions(x)
    w = oracle.World(sc)
    w.set_positions(x)
    assert sim.nodes == w.nodes == int(fem.sum()) + 8
    g, rp, cl, vl = sim.assemble()
    go, Ao = w.assemble()
    Ag = _dense(sim.nodes, rp, cl, vl)
    print(f"[abd-assemble] nodes={sim.nodes} nnzb gpu={cl.size} oracle={Ao.nnzb} "
          f"g={rel(g, go):.2e} H={rel(Ag, Ao.dense()):.2e} pairs={sim.sizes()['pairs']}",
          flush=True)
    assert sim.sizes()["pairs"] > 0
    assert rel(g, go) < 1e-10
    assert rel(Ag, Ao.dense()) < 1e-10
    e, eo = sim.energy(), w.energy()
    print(f"[abd-energy] gpu={e.tolist()} oracle={eo.tolist()}", flush=True)
    assert eo[7] > 0
    assert e[7] == pytest.approx(eo[7], rel=1e-10)
    assert e[1] == pytest.approx(eo[1], rel=1e-9)
    assert e[0] == pytest.approx(eo[0], rel=1e-10)


@pytest.mark.parametrize("precond", ["mas", "diag"])
def test_abd_newton_iterations_match(oracle, precond):
    """Newton iterates within 1e-8 with equal PCG counts, under MAS and under
    block-Jacobi (3x3 FEM node blocks, each body's 12x12 block whole:
    ABDPreconditioner)"""
    import sgipc
    sc = _scene(n_abd=2, gap=4e-4)
    if precond == "diag":
        sc.precond = sgipc.SGIPC_PRECOND_DIAG
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    sim.kappa = 1e4
    w.kappa = 1e4
    for it in range(3):
        st = sim.newton_iteration()
        so = w.newton_iteration()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[abd-newton] {precond} it={it} pcg={st['pcg_iterations']}/{so['pcg_iterations']} "
              f"alpha={st['alpha']:.6g}/{so['alpha']:.6g} err={err:.2e}", flush=True)
        assert st["converged"] == so["converged"]
        if st["converged"]:
            break
        assert st["pcg_iterations"] == so["pcg_iterations"]
        assert err < 1e-8


def test_abd_stack_trajectory(oracle):
    """two affine boxes dropped onto a soft box: 10 steps (landing and
    resting) within 1e-8 as one trajectory, the affine boxes rigid (edge
    lengths to 1e-6).  Both sides take the same Newton and PCG path: the bank
    inverses follow the same elimination order and the GPU's linear solver
    sums in a fixed order"""
    import sgipc
    sc = _scene(n_abd=2)
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    abd = sc.abd_body
    x0 = sc.x.copy()
    for step in range(10):
        sg = sim.time_step()
        so = w.time_step()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[abd-stack] step={step} newton={sg['newton_iterations']}/"
              f"{so['newton_iterations']} pcg={sg['pcg_iterations']}/{so['pcg_iterations']} "
              f"err={err:.2e} shape={sim.energy()[7]:.3e}", flush=True)
        assert sg["newton_iterations"] == so["newton_iterations"]
        assert err < 1e-8
    for b in (0, 1):
        d0 = np.linalg.norm(x0[abd == b][:, None] - x0[abd == b][None], axis=2)
        d = np.linalg.norm(xg[abd == b][:, None] - xg[abd == b][None], axis=2)
        assert np.abs(d - d0).max() < 1e-6


def test_abd_only_scene_free_fall():
    """a scene of affine bodies only (no FEM node): free fall is exact"""
    import sgipc
    from sgipc import scene as S
    x0, t0, p0 = S.hex_grid((2, 2, 2), 0.05)
    x0 = x0 - x0.mean(0) + np.array([0.0, 1.0, 0.0])
    sc = S.assemble_bodies([(x0, t0, p0, 1e9, 0.3, 1000.0, 0, True)], dt=0.01, dhat=1e-3,
                           use_ground=False)
    sc.abs_xdelta_tol = 1e-7
    sim = sgipc.Simulation(sc)
    assert sim.nodes == 4
    xs = sim.positions().copy()
    for n in range(1, 4):
        sim.time_step()
        drop = -9.8 * 0.01 ** 2 * n * (n + 1) / 2
        np.testing.assert_allclose(sim.positions(), xs + np.array([0.0, drop, 0.0]), atol=1e-10)


def _hybrid():
    import sgipc
    from sgipc import scene as S
    # 8 affine boxes (2^3 cells, kappa 1e11) fired at 2 m/s into a soft bed of
    # 8x4x8 cells (E = 5e4, bottom layer fixed), friction mu = 0.3
    sc = S.hybrid_scene(n_side=2, box_cells=2, bed_cells=(8, 4, 8), bed_size=0.4,
                        box_size=0.08, speed=2.0, gap=2e-3)
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    return sc


def test_hybrid_impact_newton_iterations_match(oracle):
    """the first Newton iterations of the impact (ABD boxes, soft bed with a
    fixed layer, friction) from the same state: iterates within 1e-8"""
    import sgipc
    sc = _hybrid()
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    for it in range(3):
        st = sim.newton_iteration()
        so = w.newton_iteration()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[hybrid-newton] it={it} pcg={st['pcg_iterations']}/{so['pcg_iterations']} "
              f"alpha={st['alpha']:.6g}/{so['alpha']:.6g} err={err:.2e}", flush=True)
        assert st["pcg_iterations"] == so["pcg_iterations"]
        assert err < 1e-8


def test_hybrid_impact_steps(oracle):
    """time steps of the impact (20-40 Newton iterations each): each step
    starts the oracle from the GPU's state.  Newton converges slowly on the
    soft bed, so when rounding (element and contact sums in another order on
    each side) splits the Newton paths, the two end points differ by more
    than the last Newton step: the steps are held to 5e-3 of the scene scale
    (1e-4 when Newton and PCG counts agree); the boxes stay rigid and the fixed
    bed layer does not move"""
    import sgipc
    sc = _hybrid()
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    abd = sc.abd_body
    fixed = sc.btype == 1
    x0 = sc.x.copy()
    for step in range(4):
        if step > 0:
            w.set_state(sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)
        sg = sim.time_step()
        so = w.time_step()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[hybrid-step] step={step} newton={sg['newton_iterations']}/"
              f"{so['newton_iterations']} pcg={sg['pcg_iterations']}/{so['pcg_iterations']} "
              f"err={err:.2e} box ymin={xg[abd >= 0, 1].min():.5f}", flush=True)
        same_path = (sg["newton_iterations"] == so["newton_iterations"]
                     and sg["pcg_iterations"] == so["pcg_iterations"])
        assert err < (1e-4 if same_path else 5e-3)
        assert np.all(np.isfinite(sim.energy()))
    np.testing.assert_array_equal(xg[fixed], x0[fixed])
    for b in range(int(abd.max()) + 1):
        d0 = np.linalg.norm(x0[abd == b][:, None] - x0[abd == b][None], axis=2)
        d = np.linalg.norm(xg[abd == b][:, None] - xg[abd == b][None], axis=2)
        assert np.abs(d - d0).max() < 1e-6


def test_hybrid_impact_trajectory_to_the_step_minimiser(oracle):
    """the same impact as one trajectory (no restarts): with PCG stopped on
    the FP64 true residual 1e-6 and Newton on |dx|_inf < 1e-6 of the scene
    diagonal, each step ends within that distance of the incremental
    potential's minimiser on both sides, whatever path rounding takes; four
    steps (70 Newton iterations each) held to 1e-5 of the scene scale"""
    import sgipc
    sc = _hybrid()
    sc.pcg_rel_tol = 1e-6
    sc.pcg_use_preconditioned_norm = 0
    sc.abs_xdelta_tol = 1e-6
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    for step in range(4):
        sg = sim.time_step()
        so = w.time_step()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[hybrid-tight] step={step} newton={sg['newton_iterations']}/"
              f"{so['newton_iterations']} pcg={sg['pcg_iterations']}/{so['pcg_iterations']} "
              f"err={err:.2e}", flush=True)
        assert sg["capped"] == 0
        assert err < 1e-5


def test_contact_table_growth(oracle, monkeypatch):
    """a contact table of 16 slots (SGIPC_CONTACT_TABLE_SLOTS) runs full on
    the first assemblies: the table grows and the assembly is redone until it
    fits; the node system still equals the oracle's (affine boxes on a soft
    box, rotated and sheared, contacts between all of them)"""
    import sgipc
    monkeypatch.setenv("SGIPC_CONTACT_TABLE_SLOTS", "16")
    sc = _scene(n_abd=2, gap=4e-4)
    sc.kappa = 1e4
    x = sc.x.copy()
    abd = sc.abd_body
    rng = np.random.default_rng(5)
    for b in (0, 1):
        c = x[abd == b].mean(0)
        A = np.eye(3) + 0.01 * rng.normal(size=(3, 3))
        x[abd == b] = c + (x[abd == b] - c) @ A.T
    sim = sgipc.Simulation(sc)
    sim.set_positions(x)
    w = oracle.World(sc)
    w.set_positions(x)
    g, rp, cl, vl = sim.assemble()
    go, Ao = w.assemble()
    Ag = _dense(sim.nodes, rp, cl, vl)
    ncoo = sim.sizes()["coo_blocks"]
    print(f"[table-growth] contact blocks={ncoo} g={rel(g, go):.2e} H={rel(Ag, Ao.dense()):.2e}",
          flush=True)
    assert ncoo > 16
    assert rel(g, go) < 1e-10
    assert rel(Ag, Ao.dense()) < 1e-10
    # a landing step with the small initial table, against the oracle
    sc2 = _scene(n_abd=2)
    sim2 = sgipc.Simulation(sc2)
    w2 = oracle.World(sc2)
    for step in range(3):
        sim2.time_step()
        w2.time_step()
    err = np.abs(sim2.positions() - w2.positions()).max() / np.abs(w2.positions()).max()
    print(f"[table-growth] 3 steps err={err:.2e}", flush=True)
    assert err < 1e-8


def test_motor_body_trajectory(oracle):
    """an affine motor box (btype 2 on its vertices) turning on a soft box:
    motor target per step, its gradient and P = strength M[6:12, 6:12] block
    on the GPU against the oracle, positions within 1e-8 per step"""
    import sgipc
    sc = _scene(n_abd=1, gap=2e-4, motor_speed=-3.0, motor_strength=200.0)
    sc.btype = sc.btype.copy()
    sc.btype[sc.abd_body == 0] = 2
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    for step in range(6):
        sg = sim.time_step()
        so = w.time_step()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[motor] step={step} newton={sg['newton_iterations']}/{so['newton_iterations']} "
              f"err={err:.2e} kin+motor={sim.energy()[1]:.4e}", flush=True)
        assert err < 1e-8
    # the box turned about its x axis
    abd = sc.abd_body
    x0, x = sc.x[abd == 0], sim.positions()[abd == 0]
    A, *_ = np.linalg.lstsq(x0 - x0.mean(0), x - x.mean(0), rcond=None)
    assert np.arctan2(A.T[2, 1], A.T[1, 1]) < -0.01


def test_contact_free_system_is_exactly_symmetric():
    """before any contact the Newton matrix comes from the element gathers
    (each lower block the exact transpose of its upper partner) and the
    affine-body node blocks (mass matrix accumulated symmetrically, packed
    shape Hessian): A equals its transpose bit for bit, FEM and affine rows
    alike"""
    import sgipc
    from sgipc import scene as S
    sc = S.hybrid_scene(n_side=2, box_cells=2, bed_cells=(8, 4, 8), bed_size=0.4,
                        box_size=0.08, speed=2.0, gap=0.05)
    sim = sgipc.Simulation(sc)
    sim.newton_iteration()
    assert int(sim.sizes()["pairs"]) == 0
    _, rowptr, cols, vals = sim.assemble()
    A = _dense(sim.nodes, rowptr, cols, vals)
    nfem = 3 * (sim.nodes - 4 * 8)
    print(f"[symmetric] n={A.shape[0]} max|A-A^T|={np.abs(A - A.T).max():.3e}", flush=True)
    assert np.abs(A[nfem:, nfem:]).max() > 0  # affine node blocks present
    assert np.array_equal(A, A.T)
    sim.close()
