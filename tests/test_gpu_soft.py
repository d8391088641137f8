"""GPU parity of the animated soft targets (GIPC.cu:6365-6435, 7527-7560) and
the animation sub-steps of a time step (IPC_Solver, GIPC.cu:11018-11060)
against the oracle: per-step positions of a bar whose end is pulled by soft
targets in two Newton loops per step (motion_rate 2), then rotated frame by
frame through sgipc_set_soft_targets; energies and the assembled system at a
shared state."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _soft_bar(k=2e3, motion_rate=2.0, shift=(0.0, 0.0, 0.01)):
    from sgipc import scene as S
    sc = S.bar_scene()
    idx = np.where(sc.x[:, 0] < 1e-9)[0]
    return S.with_soft_targets(sc, idx, sc.x[idx] + np.asarray(shift), stiffness=k,
                               motion_rate=motion_rate)


def _dense(nv, rowptr, cols, vals):
    A = np.zeros((3 * nv, 3 * nv))
    for i in range(nv):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[3 * i:3 * i + 3, 3 * cols[k]:3 * cols[k] + 3] += vals[k]
    return A


def test_soft_targets_trajectory_and_assembly(oracle):
    import sgipc
    from sgipc import scene as S
    sc = _soft_bar()
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    assert sim.animation_rate == 0.0
    origin = sc.soft_target.mean(0)
    t = sc.soft_target.copy()
    for step in range(6):
        if step >= 2:  # rotate the targets 3 degrees per frame about the bar axis
            t = S.rotate_axis_angle(t, origin, (1.0, 0.0, 0.0), np.deg2rad(3.0))
            sim.set_soft_targets(t)
            w.set_soft_targets(t)
        sg = sim.time_step()
        so = w.time_step()
        xg, xo = sim.positions(), w.positions()
        err = np.abs(xg - xo).max() / np.abs(xo).max()
        print(f"[soft] step={step} newton={sg['newton_iterations']}/{so['newton_iterations']} "
              f"pcg={sg['pcg_iterations']}/{so['pcg_iterations']} err={err:.2e} "
              f"soft={sim.energy()[8]:.4e}", flush=True)
        assert sg["newton_iterations"] == so["newton_iterations"] >= 2
        assert sim.animation_rate == w.anim_rate == 1.0
        assert err < 1e-8
    # energies and the assembled system at the GPU's state
    w.set_state(sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)
    e, eo = sim.energy(), w.energy()
    assert eo[8] > 0
    assert rel(e, eo) < 1e-12 and abs(e[8] - eo[8]) <= 1e-12 * eo[8]
    g, rp, cl, vl = sim.assemble()
    go, Ao = w.assemble()
    assert rel(g, go) < 1e-10
    assert rel(_dense(sc.n_verts, rp, cl, vl), Ao.dense()) < 1e-10


def test_soft_target_validation():
    import sgipc
    from dataclasses import replace
    sc = _soft_bar()
    bad = replace(sc, soft_index=np.array([0, 0], np.uint32),
                  soft_target=np.zeros((2, 3)))
    with pytest.raises(sgipc.SgipcError, match="two soft targets"):
        sgipc.Simulation(bad)
    fixed = sc.btype.copy()
    fixed[sc.soft_index[0]] = 1
    with pytest.raises(sgipc.SgipcError, match="fixed"):
        sgipc.Simulation(replace(sc, btype=fixed))
