"""Spatial re-sorting keys (GIPC.cu:142-198 hash_code, :279-341
_calcTetMChash, :412-461 _calcVertMChash): the oracle's restatement pinned by
known answers, an independent bit-by-bit Morton interleave, the octree
property of Morton order on a dyadic grid, and the host-side scene
permutation (sgipc.sort) driven by oracle orders.  The reference has no tests
of these utilities (no caller, no fixture)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "stiff-gipc_amd"))
import pyoracle as po  # noqa: E402


def _interleave(X, Y, Z):
    """Morton code by bit loop: bit b of X at 3b+2, of Y at 3b+1, of Z at 3b"""
    c = 0
    for b in range(10):
        c |= ((X >> b) & 1) << (3 * b + 2)
        c |= ((Y >> b) & 1) << (3 * b + 1)
        c |= ((Z >> b) & 1) << (3 * b)
    return c


def test_morton_known_answers():
    h = po.lib().orc_hash_code
    assert h(-1, 0.0, 0.0, 0.0) == 0
    assert h(-1, 0.5, 0.0, 0.0) == 1 << 29          # X = 512: bit 9 -> bit 29
    assert h(-1, 0.0, 0.5, 0.0) == 1 << 28
    assert h(-1, 0.0, 0.0, 0.5) == 1 << 27
    assert h(-1, 1.0, 1.0, 1.0) == (1 << 30) - 1    # clamped to cell 1023
    assert h(-1, 7.0, -3.0, 1.0 / 1024) == _interleave(1023, 0, 1)
    assert h(-1, float("nan"), 0.0, 0.0) == 0        # flat axis: cell 0


def test_morton_equals_bit_loop_interleave():
    rng = np.random.default_rng(3)
    for X, Y, Z in rng.integers(0, 1024, (400, 3)):
        # the cell centre maps back to the same cell
        got = po.lib().orc_hash_code(-1, (X + 0.5) / 1024, (Y + 0.5) / 1024, (Z + 0.5) / 1024)
        assert got == _interleave(int(X), int(Y), int(Z))


@pytest.mark.parametrize("typ,expect", [
    (0, (3 * 1024 + 2) * 1024 + 1),   # z, y, x from most to least significant
    (1, (2 * 1024 + 3) * 1024 + 1),   # y, z, x
    (2, (1 * 1024 + 3) * 1024 + 2),   # x, z, y
    (3, (3 * 1024 + 1) * 1024 + 2),   # z, x, y
    (4, (2 * 1024 + 1) * 1024 + 3),   # y, x, z
    (5, (1 * 1024 + 2) * 1024 + 3),   # x, y, z
])
def test_lexicographic_known_answers(typ, expect):
    """cells (X, Y, Z) = (1, 2, 3)"""
    assert po.lib().orc_hash_code(typ, 1.5 / 1024, 2.5 / 1024, 3.5 / 1024) == expect


def test_tet_type_from_extents():
    t = po.lib().orc_tet_hash_type
    # the longest axis varies fastest in the key; ties fall to type 5
    assert t(3.0, 2.0, 1.0) == 0
    assert t(3.0, 1.0, 2.0) == 1
    assert t(1.0, 3.0, 2.0) == 2
    assert t(2.0, 3.0, 1.0) == 3
    assert t(2.0, 1.0, 3.0) == 4
    assert t(1.0, 2.0, 3.0) == 5
    assert t(1.0, 1.0, 1.0) == 5
    assert t(2.0, 2.0, 1.0) == 5


def test_vertex_keys_sorted_permutation_and_codes():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2000, 3)) * [3.0, 1.0, 0.5]
    keys, order = po.spatial_keys(x)
    assert np.all(np.diff(keys.astype(np.uint64)) > 0)
    assert np.array_equal(np.sort(order), np.arange(2000))
    assert np.array_equal(keys & np.uint64(0xFFFFFFFF), order.astype(np.uint64))
    lo, hi = x.min(0), x.max(0)
    cells = np.minimum(np.maximum(np.floor((x - lo) / (hi - lo) * 1024), 0), 1023).astype(int)
    for i in rng.integers(0, 2000, 50):
        j = int(np.flatnonzero(order == i)[0])
        assert int(keys[j]) >> 32 == _interleave(*cells[i])


def test_morton_order_of_a_dyadic_grid_is_octree_order():
    """points i/16 (i < 16 per axis) plus the far corner: each run of 8 in
    the sorted order is a 2x2x2 cube and each run of 64 a 4x4x4 cube"""
    g = np.stack(np.meshgrid(*[np.arange(16)] * 3, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(1)
    g = g[rng.permutation(len(g))]
    x = np.concatenate([g / 16.0, [[1.0, 1.0, 1.0]]])
    _, order = po.spatial_keys(x)
    assert order[-1] == len(g)
    s = g[order[:-1]]
    for run, side in ((8, 2), (64, 4), (512, 8)):
        blocks = s.reshape(-1, run, 3) // side
        assert np.all(blocks == blocks[:, :1]), run


def test_ties_keep_index_order_and_flat_axis():
    x = np.zeros((37, 3))
    x[:, 0] = np.repeat([0.0, 1.0], [20, 17])
    keys, order = po.spatial_keys(x)      # y, z flat: cells 0
    assert np.array_equal(order, np.arange(37))
    assert int(keys[-1]) >> 32 == _interleave(1023, 0, 0)


def test_tet_keys_follow_the_centroid():
    from sgipc.scene import bar_scene
    sc = bar_scene()
    x = sc.x * [1.0, 0.8, 0.5]             # extents 0.8 > 0.16 > 0.1: type 0
    keys, order = po.spatial_keys(x, sc.tets)
    assert np.array_equal(np.sort(order), np.arange(sc.n_tets))
    lo, hi = x.min(0), x.max(0)
    ext = hi - lo
    assert po.lib().orc_tet_hash_type(*ext) == 0
    v = x[sc.tets.astype(np.int64)]
    c = ((v[:, 0] + v[:, 1]) + (v[:, 2] + v[:, 3])) * 0.25   # the reference's centroid
    # type 0 packs (z, y, x) from the most significant field down: the sorted
    # centroids' (z, y, x) cells ascend lexicographically
    cz, cy, cx = [np.floor((c[order, k] - lo[k]) / ext[k] * 1024).astype(np.int64)
                  for k in (2, 1, 0)]
    lex = (cz * 1024 + cy) * 1024 + cx
    assert np.all(np.diff(lex) >= 0)
    assert np.array_equal(keys >> np.uint64(32), lex.astype(np.uint64))


def _scene_invariants(sc, r):
    from sgipc.sort import _n_free
    s = r.scene
    n = sc.n_verts
    assert np.array_equal(s.x, sc.x[r.vert_order])
    assert np.array_equal(r.vert_map[r.vert_order], np.arange(n))
    nf = _n_free(sc)
    assert np.array_equal(r.vert_order[nf:], np.arange(nf, n))
    # every element keeps its vertex positions
    for old, new in ((sc.faces, s.faces), (sc.edges, s.edges), (sc.tris, s.tris)):
        if old is not None:
            assert np.array_equal(sc.x[old.astype(np.int64)], s.x[new.astype(np.int64)])
    assert np.array_equal(sc.x[sc.tets[r.tet_order].astype(np.int64)],
                          s.x[s.tets.astype(np.int64)])
    assert np.array_equal(s.E, sc.E[r.tet_order])
    assert np.array_equal(np.sort(sc.x[sc.surf_verts.astype(np.int64)], 0),
                          np.sort(s.x[s.surf_verts.astype(np.int64)], 0))
    assert np.array_equal(s.btype, sc.btype[r.vert_order])
    assert np.array_equal(s.body_id, sc.body_id[r.vert_order])
    p = s.part_id
    assert np.all(np.diff(p) >= 0)
    assert np.bincount(p[:nf]).max() <= 16
    if sc.soft_index is not None:
        assert np.array_equal(s.x[s.soft_index.astype(np.int64)],
                              sc.x[sc.soft_index.astype(np.int64)])


@pytest.mark.parametrize("name", ["bar", "abd", "cloth", "soft"])
def test_scene_permutation_with_oracle_orders(name):
    from sgipc import scene as S
    from sgipc.sort import _n_free, spatial_sort
    sc = {"bar": lambda: S.bar_scene(),
          "abd": lambda: S.abd_scene(n_abd=2),
          "cloth": lambda: S.cloth_scene(n=8),
          "soft": lambda: S.with_soft_targets(S.bar_scene(), np.arange(0, 40, 3))}[name]()
    nf = _n_free(sc)
    _, vo = po.spatial_keys(sc.x[:nf])
    # tet keys of the re-indexed tets on the re-ordered positions
    order = np.concatenate([vo.astype(np.int64), np.arange(nf, sc.n_verts)])
    vmap = np.empty(sc.n_verts, np.int64)
    vmap[order] = np.arange(sc.n_verts)
    _, to = po.spatial_keys(sc.x[order], vmap[sc.tets.astype(np.int64)])
    r = spatial_sort(sc, orders=(vo, to))
    _scene_invariants(sc, r)
    assert r.scene.meta["spatially_sorted"]


def test_bad_orders_are_refused():
    from sgipc import scene as S
    from sgipc.sort import spatial_sort
    sc = S.bar_scene()
    with pytest.raises(ValueError):
        spatial_sort(sc, orders=(np.zeros(sc.n_verts, np.int64), np.arange(sc.n_tets)))
    with pytest.raises(ValueError):
        spatial_sort(sc, bank=17)
