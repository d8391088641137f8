"""Newton-step level properties of the oracle (no GPU)."""
import numpy as np
import pytest


def test_free_fall_step_is_exact_translation(oracle):
    from sgipc import scene as S
    sc = S.bar_scene(drop=1.0)  # far from the ground: no contacts
    w = oracle.World(sc)
    xt = w.xtilde()
    w.time_step()
    # x_tilde is a rigid translation, F = I there, so it is the minimiser
    assert np.abs(w.positions() - xt).max() < 1e-9


def test_newton_line_search_decreases_energy(oracle):
    from sgipc import scene as S
    sc = S.stack_scene(n=2, cells=3)
    w = oracle.World(sc)
    w.kappa = 1e5
    rng = np.random.default_rng(0)
    w.set_xtilde(sc.x + 5e-4 * rng.normal(size=sc.x.shape))
    for _ in range(4):
        e0 = w.energy()[0]
        st = w.newton_iteration()
        if st["converged"]:
            break
        assert w.energy()[0] <= e0 * (1 + 1e-12)


def test_bar_drop_stays_above_ground(oracle):
    from sgipc import scene as S
    sc = S.bar_scene()
    w = oracle.World(sc)
    ymin = []
    for _ in range(20):
        w.time_step()
        ymin.append(w.positions()[:, 1].min())
    assert min(ymin) > 0.0
    assert ymin[-1] < 1e-3  # it has reached the ground band (dhat = 1e-3)
