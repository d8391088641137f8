"""GPU spatial re-sorting keys (GIPC.cu:279-461) against the oracle, bit for
bit (codes, tie order, permutation), from tiny inputs to the 983,040-tet bed
of the hybrid impact scene; and a Morton-sorted scene (sgipc.spatial_sort)
that is the same physical system: energies, gradient and Hessian at a shared
state agree under the permutation, and a short trajectory follows the
unsorted one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _clouds():
    rng = np.random.default_rng(11)
    flat = rng.random((3000, 3))
    flat[:, 1] = 0.25                                   # a flat axis: cell 0
    dup = np.repeat(rng.random((50, 3)), 7, axis=0)     # 7-fold ties
    return {
        "one": np.array([[0.3, -1.0, 2.0]]),
        "seven": rng.random((7, 3)),
        "cloud1k": rng.normal(size=(1000, 3)) * [5.0, 1.0, 0.2],
        "cloud100k": rng.random((100003, 3)) * [1.0, 2.0, 3.0] - 7.0,
        "flat": flat,
        "ties": dup[rng.permutation(len(dup))],
    }


@pytest.mark.parametrize("name", list(_clouds()))
def test_vertex_keys_bit_exact(oracle, name):
    import sgipc
    x = _clouds()[name]
    print(f"vertex keys {name} n={len(x)}", flush=True)
    k, o = sgipc.spatial_sort_keys(x)
    ko, oo = oracle.spatial_keys(x)
    assert np.array_equal(k, ko)
    assert np.array_equal(o, oo)


@pytest.mark.parametrize("name", ["bar", "cubes", "hybrid"])
def test_scene_keys_bit_exact(oracle, name):
    """vertex and tet keys of the benchmark scenes (hybrid: 983,040 bed tets
    plus 512 boxes of 1,296 tets, full size)"""
    import sgipc
    from sgipc import scene as S
    sc = {"bar": S.bar_scene, "cubes": lambda: S.cubes_scene(n_side=3, cells=4),
          "hybrid": S.hybrid_scene}[name]()
    print(f"scene keys {name} nv={sc.n_verts} nt={sc.n_tets}", flush=True)
    k, o = sgipc.spatial_sort_keys(sc.x)
    ko, oo = oracle.spatial_keys(sc.x)
    assert np.array_equal(k, ko) and np.array_equal(o, oo)
    k, o = sgipc.spatial_sort_keys(sc.x, sc.tets)
    ko, oo = oracle.spatial_keys(sc.x, sc.tets)
    assert np.array_equal(k, ko) and np.array_equal(o, oo)


def test_refuses_bad_input():
    import sgipc
    from sgipc import lib
    with pytest.raises(sgipc.SgipcError):
        sgipc.spatial_sort_keys(np.zeros((4, 3)), np.array([[0, 1, 2, 9]], np.uint32))
    assert b"out of range" in lib().sgipc_last_error()
    with pytest.raises(sgipc.SgipcError):
        sgipc.spatial_sort_keys(np.zeros((0, 3)))


def _dense(nn, rowptr, cols, vals):
    A = np.zeros((3 * nn, 3 * nn))
    for i in range(nn):
        for k in range(rowptr[i], rowptr[i + 1]):
            A[3 * i:3 * i + 3, 3 * cols[k]:3 * cols[k] + 3] += vals[k]
    return A


def _dof(vmap):
    return (3 * vmap[:, None] + np.arange(3)).ravel()


@pytest.mark.parametrize("name", ["bar", "stack"])
def test_sorted_scene_is_the_same_system(name):
    import sgipc
    from sgipc import scene as S
    sc = {"bar": S.bar_scene, "stack": lambda: S.stack_scene(n=2, cells=3)}[name]()
    r = sgipc.spatial_sort(sc)
    assert r.scene.part_id.max() < sc.part_id.max() + 16
    a, b = sgipc.Simulation(sc), sgipc.Simulation(r.scene)
    for s in (a, b):
        s.set_pcg(rel_tol=1e-10)
    for _ in range(3):
        a.time_step()
    x, v, xt, kap = a.positions(), a.velocities(), a.xtilde(), a.kappa
    o = r.vert_order
    b.set_state(x[o], v[o], xt[o], kap)
    a.set_state(x, v, xt, kap)
    ea, eb = a.energy(), b.energy()
    assert np.allclose(ea, eb, rtol=1e-11, atol=1e-14 * abs(ea[0]))
    ga, rpa, ca, va = a.assemble()
    gb, rpb, cb, vb = b.assemble()
    d = _dof(r.vert_map)
    assert np.allclose(gb[d], ga, rtol=1e-11, atol=1e-12 * np.abs(ga).max())
    nn = a.nodes
    Aa, Ab = _dense(nn, rpa, ca, va), _dense(nn, rpb, cb, vb)
    assert np.allclose(Ab[np.ix_(d, d)], Aa, rtol=1e-11, atol=1e-12 * np.abs(Aa).max())


def test_sorted_trajectory_follows_the_unsorted_one():
    """three steps of a bar dropped on the ground (FP64 MAS, tight PCG): the
    two orderings differ only in the PCG's banks and summation order"""
    import sgipc
    from sgipc import scene as S
    sc = S.bar_scene()
    sc.mas_precision = sgipc.SGIPC_MAS_FP64
    r = sgipc.spatial_sort(sc)
    a, b = sgipc.Simulation(sc), sgipc.Simulation(r.scene)
    for s in (a, b):
        s.set_pcg(rel_tol=1e-10)
    diag = np.linalg.norm(sc.x.max(0) - sc.x.min(0))
    for step in range(3):
        a.time_step()
        b.time_step()
        err = np.abs(b.positions()[r.vert_map] - a.positions()).max()
        print(f"step {step}: max |dx| = {err:.3e} (diag {diag:.3f})", flush=True)
        assert err < 1e-7 * diag
