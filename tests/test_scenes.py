"""Host-side checks of the benchmark scene builders (no GPU): the sizes of
the configurations they stand for, body ordering (FEM vertices before
affine-body vertices, as the node system expects), non-overlapping bodies,
velocities and fixed layers."""
import numpy as np


def test_cubes_scene_sizes():
    from sgipc import scene as S
    sc = S.cubes_scene(n_side=4, cells=8)
    assert sc.n_tets == 64 * 3072 and sc.n_verts == 64 * 729
    assert np.all(np.diff(sc.part_id) >= 0)
    assert sc.x[:, 1].min() > 0.5  # dropped from 0.5 m


def test_hybrid_scene_layout():
    from sgipc import scene as S
    sc = S.hybrid_scene(n_side=2, box_cells=2, bed_cells=(8, 4, 8), bed_size=0.4, box_size=0.08)
    abd = sc.abd_body
    nbed = 9 * 5 * 9
    assert sc.n_tets == 8 * 4 * 8 * 6 + 8 * 8 * 6
    assert np.all(abd[:nbed] == -1) and np.all(abd[nbed:] >= 0)
    assert np.all(np.diff(abd[nbed:]) >= 0)                   # bodies contiguous
    assert np.all(sc.btype[:nbed][sc.x[:nbed, 1] <= 1e-12] == 1)
    assert np.all(sc.btype[nbed:] == 0)
    np.testing.assert_allclose(sc.v[nbed:], np.tile([0.0, -10.0, 0.0], (sc.n_verts - nbed, 1)))
    np.testing.assert_allclose(sc.v[:nbed], 0.0)
    top = sc.x[:nbed, 1].max()
    assert sc.x[nbed:, 1].min() > top                          # above the bed
    # boxes do not overlap: their bounding spheres are apart
    ctr = np.array([sc.x[abd == b].mean(0) for b in range(8)])
    rad = np.array([np.linalg.norm(sc.x[abd == b] - ctr[b], axis=1).max() for b in range(8)])
    d = np.linalg.norm(ctr[:, None] - ctr[None], axis=2) + np.eye(8) * 1e9
    assert np.all(d > rad[:, None] + rad[None])


def test_hybrid_scene_full_size():
    from sgipc import scene as S
    sc = S.hybrid_scene()
    assert sc.n_tets == 64 * 64 * 40 * 6 + 512 * 1296
    assert int(sc.abd_body.max()) + 1 == 512
    nodes = int((sc.abd_body < 0).sum()) + 4 * 512
    assert nodes == 65 * 65 * 41 + 2048


def test_cloth_scene_boxes_apart():
    from sgipc import scene as S
    sc = S.cloth_scene(n=32, size=1.0, height=0.3, n_boxes=16, box_size=0.1, box_cells=2,
                       abd_boxes=True, abd_kappa=1e11)
    abd = sc.abd_body
    assert int(abd.max()) + 1 == 16
    ctr = np.array([sc.x[abd == b].mean(0) for b in range(16)])
    dx = np.abs(ctr[:, None, 0] - ctr[None, :, 0])
    dz = np.abs(ctr[:, None, 2] - ctr[None, :, 2])
    sep = np.maximum(dx, dz) + np.eye(16) * 9
    assert np.all(sep >= 0.1 + 0.02 - 1e-12)
    assert sc.tris.shape[0] == 2 * 31 * 31
