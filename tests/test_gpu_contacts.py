"""GPU parity of the broad phase and the step-size bounds at benchmark sizes:
bit-exact LBVH (keys, nodes, AABBs) and exact contact-pair multisets on the
full 64-cube surface (49,152 faces, 73,728 edges) with cube-cube and ground
contacts, and on an impact state of affine boxes in a soft bed; the Newton
step bounds (ground, ACCD self bound, CFL, full-CCD candidate count and the
resulting alpha) and the full-CCD candidate multisets against the oracle."""
import numpy as np
import pytest

from helpers import bring_into_contact, lower_onto_ground

pytestmark = pytest.mark.gpu


def _key(p):
    return sorted(map(tuple, np.asarray(p).tolist()))


def _check_lbvh_and_pairs(oracle, sim, sc, x, tag):
    for which, prims in ((0, sc.faces), (1, sc.edges)):
        k, nd, bv = sim.lbvh(which)
        ko, ndo, bvo = oracle.lbvh(x, prims)
        print(f"[{tag}] lbvh which={which} n={prims.shape[0]}", flush=True)
        np.testing.assert_array_equal(k, ko)
        np.testing.assert_array_equal(nd, ndo)
        np.testing.assert_array_equal(bv.view(np.uint64), bvo.view(np.uint64))
    pairs, counts, ground = sim.contacts()
    po, co = oracle.contact_pairs(sc, x, "bvh")
    print(f"[{tag}] contacts gpu={counts.tolist()} oracle={co.tolist()} ground={ground.size}",
          flush=True)
    assert counts[0] > 0
    assert _key(pairs) == _key(po)
    np.testing.assert_array_equal(counts, co)
    return pairs, ground


def _cubes64():
    from sgipc import scene as S
    sc = S.cubes_scene(n_side=4, cells=8, size=0.2, spacing=0.6, fall=0.5, seed=0)
    sc.x = bring_into_contact(sc, 0, 1)
    sc.x = bring_into_contact(sc, 5, 21)
    sc.x = lower_onto_ground(sc.x, sc.dhat)
    return sc


def test_full_64_cube_surface(oracle):
    import sgipc
    sc = _cubes64()
    sim = sgipc.Simulation(sc)
    x = sim.positions()
    _, ground = _check_lbvh_and_pairs(oracle, sim, sc, x, "cubes64")
    d = x[:, 1] ** 2
    expect = np.sort(sc.surf_verts[d[sc.surf_verts] <= sc.dhat ** 2])
    assert expect.size > 0
    np.testing.assert_array_equal(np.sort(ground), expect)


def _impact():
    from sgipc import scene as S
    # 64 affine boxes (3^3 cells) at 4 m/s into a 16x8x16-cell bed
    return S.hybrid_scene(n_side=4, box_cells=3, bed_cells=(16, 8, 16), bed_size=0.8,
                          box_size=0.1, speed=4.0, gap=2e-3)


def test_impact_state_lbvh_and_pairs(oracle):
    import sgipc
    sc = _impact()
    sim = sgipc.Simulation(sc)
    for _ in range(2):
        sim.time_step()
    x = sim.positions()
    _check_lbvh_and_pairs(oracle, sim, sc, x, "impact")


def _bounds_close(g, o, tag):
    print(f"[{tag}] gpu={g} oracle={o}", flush=True)
    assert g["n_candidates"] == o["n_candidates"]
    for k in ("alpha_ground", "alpha_self", "max_speed", "alpha_cfl", "alpha"):
        assert abs(g[k] - o[k]) <= 1e-13 * max(abs(o[k]), 1e-300), k


def test_ccd_bound_bar_and_stack(oracle):
    import sgipc
    from sgipc import scene as S
    sc = S.bar_scene()
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    rng = np.random.default_rng(1)
    for s in (1.0, 0.05):
        dx = np.ones((sc.n_verts, 1)) * np.array([0.2, s, -0.1])
        dx += 0.01 * rng.normal(size=dx.shape)
        _bounds_close(sim.ccd_bound(dx), w.ccd_bound(dx), f"bar s={s}")
    sc = S.stack_scene(n=2, cells=3)
    sim = sgipc.Simulation(sc)
    w = oracle.World(sc)
    top = sc.x[:, 1] > sc.x[:, 1].mean()
    for u in (0.01, 0.002):
        dx = np.zeros_like(sc.x)
        dx[top, 1] = u
        dx[top] += 1e-3 * rng.normal(size=(int(top.sum()), 3))
        g, o = sim.ccd_bound(dx), w.ccd_bound(dx)
        _bounds_close(g, o, f"stack u={u}")
        assert g["n_candidates"] > 0
        cg = sim.ccd_candidates(dx, g["alpha"])
        co = w.ccd_candidates(dx, o["alpha"])
        assert len(cg) > 0 and _key(cg) == _key(co)


def test_ccd_bound_impact(oracle):
    """mid-impact state of the affine boxes, direction towards the inertial
    target (x - x_tilde): ACCD- and CFL-limited"""
    import sgipc
    sc = _impact()
    sim = sgipc.Simulation(sc)
    for _ in range(2):
        sim.time_step()
    w = oracle.World(sc)
    w.set_state(sim.positions(), sim.velocities(), sim.xtilde(), sim.kappa)
    # set_state refits the affine bodies and re-projects their vertices (two
    # implementations, rounding apart); the bounds are compared at the GPU's
    # positions exactly
    w.set_positions_exact(sim.positions())
    np.testing.assert_array_equal(w.positions(), sim.positions())
    dx = sim.positions() - sim.xtilde()
    g, o = sim.ccd_bound(dx), w.ccd_bound(dx)
    _bounds_close(g, o, "impact")
    assert g["n_candidates"] > 0 and g["alpha"] < 1.0
    for a in (g["alpha"], 1.0):
        cg = sim.ccd_candidates(dx, a)
        co = w.ccd_candidates(dx, a)
        print(f"[impact] candidates alpha={a:.4g}: {len(cg)}", flush=True)
        assert len(cg) > 0 and _key(cg) == _key(co)
