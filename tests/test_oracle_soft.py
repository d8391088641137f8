"""Oracle pins of the animated soft targets (GIPC.cu:6365-6435, 7527-7560) and
the animation sub-steps of a time step (IPC_Solver, GIPC.cu:11018-11060):
energy = the closed form 0.5 d |r (x - t)|^2, its gradient d r^2 (x - t) and
Hessian r^2 d I as the difference of two worlds that differ only in the
stiffness, the assembled gradient = central differences of the total energy,
the rate ramp r = 1/m, ..., 1 with one Newton loop per rate, and targets that
pull their vertices."""
import numpy as np
import pytest


def _soft_bar(k=2e3, motion_rate=1.0, shift=(0.0, 0.0, 0.01), gravity=(0.0, -9.8, 0.0)):
    from sgipc import scene as S
    sc = S.bar_scene(gravity=gravity)
    idx = np.where(sc.x[:, 0] < 1e-9)[0]          # the x = 0 end face
    t = sc.x[idx] + np.asarray(shift)
    return S.with_soft_targets(sc, idx, t, stiffness=k, motion_rate=motion_rate)


def test_soft_energy_gradient_hessian_closed_form(oracle):
    sc = _soft_bar()
    w1 = oracle.World(sc)
    from dataclasses import replace
    w0 = oracle.World(replace(sc, soft_stiffness=0.0))
    r = 0.7
    w1.anim_rate = r
    w0.anim_rate = r
    x = w1.positions()
    e1, e0 = w1.energy(), w0.energy()
    ref = 0.0
    for i, v in enumerate(sc.soft_index):
        d = (x[v] - sc.soft_target[i]) * r
        ref += sc.soft_stiffness * (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) * 0.5
    assert e1[8] == ref and e0[8] == 0.0
    assert e1[0] == pytest.approx(e0[0] + ref, rel=1e-14)
    g1, A1 = w1.assemble()
    g0, A0 = w0.assemble()
    dg = (g1 - g0).reshape(-1, 3)
    expect = np.zeros_like(dg)
    expect[sc.soft_index] = sc.soft_stiffness * r * r * (x[sc.soft_index] - sc.soft_target)
    scale = np.abs(g1).max()
    assert np.abs(dg - expect).max() <= 1e-14 * scale
    D1, D0 = A1.dense(), A0.dense()
    dH = D1 - D0
    want = np.zeros_like(dH)
    for v in sc.soft_index:
        want[3 * v:3 * v + 3, 3 * v:3 * v + 3] = r * r * sc.soft_stiffness * np.eye(3)
    assert np.abs(dH - want).max() <= 1e-14 * np.abs(D1).max()


def test_soft_gradient_is_energy_derivative(oracle):
    sc = _soft_bar(k=5e3)
    w = oracle.World(sc)
    w.time_step()          # x~, kappa and the rate (1 after the step) in place
    assert w.anim_rate == 1.0
    x = w.positions()
    g, _ = w.assemble()
    rng = np.random.default_rng(0)
    verts = np.concatenate([sc.soft_index[:4], rng.integers(0, sc.n_verts, 4)])
    h = 1e-6
    for v in verts:
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[v, c] += h
            xm[v, c] -= h
            w.set_positions(xp)
            ep = w.energy()[0]
            w.set_positions(xm)
            em = w.energy()[0]
            fd = (ep - em) / (2 * h)
            assert fd == pytest.approx(g[3 * v + c], rel=1e-5, abs=1e-7 * np.abs(g).max())
    w.set_positions(x)


def test_rate_ramp_runs_one_newton_loop_per_rate(oracle):
    sc4 = _soft_bar(motion_rate=4.0)
    w4 = oracle.World(sc4)
    assert w4.anim_rate == 0.0            # GIPC.cuh:114 before the first step
    st = w4.time_step()
    assert w4.anim_rate == 1.0            # 0.25 + 0.25 + 0.25 + 0.25
    w1 = oracle.World(_soft_bar(motion_rate=1.0))
    s1 = w1.time_step()
    # each of the 4 loops solves at least once (its first iteration has no stop test)
    assert st["newton_iterations"] >= 4 and st["newton_iterations"] >= s1["newton_iterations"]


def test_soft_targets_pull_their_vertices(oracle):
    from sgipc import scene as S
    sc = _soft_bar(k=2e4, shift=(0.0, 0.0, 0.0), gravity=(0.0, 0.0, 0.0))
    sc.use_ground = False
    w = oracle.World(sc)
    # rotate the targets about the bar's long axis by 10 degrees, in 5 frames
    origin = sc.x[sc.soft_index].mean(0)
    t = sc.x[sc.soft_index].copy()
    for _ in range(5):
        t = S.rotate_axis_angle(t, origin, (1.0, 0.0, 0.0), np.deg2rad(2.0))
        w.set_soft_targets(t)
        w.time_step()
    x = w.positions()[sc.soft_index]
    err0 = np.linalg.norm(sc.x[sc.soft_index] - t, axis=1).max()
    err = np.linalg.norm(x - t, axis=1).max()
    assert err < 0.25 * err0
