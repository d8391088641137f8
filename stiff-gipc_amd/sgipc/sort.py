"""Spatial re-sorting of a scene (the reference's re-ordering utilities,
StiffGIPC/GIPC.cu: _calcVertMChash :412-461, _calcTetMChash :279-341,
_updateVertexes :369-389, _updateTetrahedras :391-410, _updateTopology
:343-367, updateSurfaces / updateSurfaceEdges / updateSurfaceVerts
:9578-9617; its sortGeometry :9735-9741 is an empty placeholder).

spatial_sort(scene) returns a copy whose elastic and cloth vertices (every
vertex before the affine bodies) are in Morton order and whose tets are in
centroid-key order, with every index array remapped, and whose MAS banks are
runs of 16 consecutive vertices of the new order.  The keys and their sort
run on the GPU (sgipc_spatial_sort_keys).  Affine-body vertices keep their
place: the system numbers them after the elastic vertices, grouped by body.
"""
from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from .scene import Scene


@dataclass
class SortResult:
    scene: Scene
    vert_order: np.ndarray    # (N,) new vertex i was old vertex vert_order[i]
    vert_map: np.ndarray      # (N,) old vertex j is new vertex vert_map[j]
    tet_order: np.ndarray     # (T,) new tet e was old tet tet_order[e]


def _n_free(sc: Scene) -> int:
    """vertices before the first affine-body vertex"""
    if sc.abd_body is None:
        return sc.n_verts
    a = np.flatnonzero(np.asarray(sc.abd_body) >= 0)
    return int(a[0]) if a.size else sc.n_verts


def spatial_sort(sc: Scene, bank: int = 16, orders=None) -> SortResult:
    """Morton-ordered copy of sc.  bank: vertices per MAS bank of the new
    part_id (the level-0 bank size, at most 16).  orders: (vertex order of
    the elastic vertices, tet order) to apply instead of computing them on
    the GPU."""
    if not 1 <= bank <= 16:
        raise ValueError("bank must be in 1..16")
    nf = _n_free(sc)
    n = sc.n_verts
    if orders is None:
        from ._lib import spatial_sort_keys
        _, vo = spatial_sort_keys(sc.x[:nf]) if nf else (None, np.zeros(0, np.uint32))
    else:
        vo = np.asarray(orders[0])
    vo = np.asarray(vo, np.int64)
    if vo.shape != (nf,) or (nf and not np.array_equal(np.sort(vo), np.arange(nf))):
        raise ValueError("vertex order is not a permutation of the elastic vertices")
    order = np.concatenate([vo, np.arange(nf, n)])
    vmap = np.empty(n, np.int64)
    vmap[order] = np.arange(n)

    def remap(a):
        return None if a is None else vmap[np.asarray(a, np.int64)].astype(np.uint32)

    x = sc.x[order].copy()
    tets = remap(sc.tets)
    if sc.n_tets:
        if orders is None:
            from ._lib import spatial_sort_keys
            _, to = spatial_sort_keys(x, tets)
        else:
            to = np.asarray(orders[1])
        to = np.asarray(to, np.int64)
        if not np.array_equal(np.sort(to), np.arange(sc.n_tets)):
            raise ValueError("tet order is not a permutation of the tets")
    else:
        to = np.zeros(0, np.int64)
    part = np.asarray(sc.part_id, np.int64)
    new_part = np.empty(n, np.int32)
    new_part[:nf] = np.arange(nf) // bank
    top = (nf + bank - 1) // bank
    if n > nf:
        # affine-body vertices: their ids in the same order after the free banks
        tail = part[nf:]
        new_part[nf:] = top + (tail - tail.min())
    out = replace(
        sc, x=x, tets=tets[to] if sc.n_tets else tets,
        E=np.asarray(sc.E)[to] if sc.n_tets else sc.E,
        nu=np.asarray(sc.nu)[to] if sc.n_tets else sc.nu,
        rho=np.asarray(sc.rho)[to] if sc.n_tets else sc.rho,
        part_id=new_part, body_id=np.asarray(sc.body_id)[order].copy(),
        btype=np.asarray(sc.btype)[order].copy(),
        faces=remap(sc.faces), edges=remap(sc.edges), surf_verts=remap(sc.surf_verts),
        v=None if sc.v is None else np.asarray(sc.v)[order].copy(),
        x_rest=None if sc.x_rest is None else np.asarray(sc.x_rest)[order].copy(),
        tris=remap(sc.tris),
        abd_body=None if sc.abd_body is None else np.asarray(sc.abd_body)[order].copy(),
        soft_index=remap(sc.soft_index),
        meta=dict(sc.meta, spatially_sorted=True))
    return SortResult(out, order, vmap, to)
