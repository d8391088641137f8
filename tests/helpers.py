"""Shared test helpers (scene construction and comparisons)."""
import numpy as np


def assembled_system(scene, perturb=0.0, seed=0, steps=0):
    """Oracle world for the scene; optionally perturb positions so the elastic
    Hessian is away from rest.  Returns (world, grad, bsr)."""
    import pyoracle as O
    w = O.World(scene)
    if perturb:
        rng = np.random.default_rng(seed)
        x = scene.x + perturb * rng.normal(size=scene.x.shape)
        w.set_positions(x)
    for _ in range(steps):
        w.time_step()
    g, A = w.assemble()
    return w, g, A


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def deep_hierarchy_system(nblob=4096, seed=0):
    """Synthetic SPD block system whose MAS hierarchy has two large coarse
    levels followed by small ones: 16-vertex chains ("blobs", one bank each);
    inside each level-1 bank (16 blobs) blobs are joined in pairs, giving 8
    level-2 nodes per level-1 bank; pairs of consecutive level-1 banks are
    joined in a zigzag, so each level-2 bank is connected and level 3 has one
    node per level-2 bank.  With 4096 blobs: level 1 = 256 banks, level 2 =
    128 banks, level 3 = 8 banks.  Returns (bsr, part_id, rhs)."""
    import pyoracle as O
    rng = np.random.default_rng(seed)
    nv = 16 * nblob
    edges = []
    for b in range(nblob):
        v0 = 16 * b
        edges += [(v0 + k, v0 + k + 1) for k in range(15)]
    first = lambda blob: 16 * blob
    for b1 in range(nblob // 16):
        for q in range(0, 16, 2):
            edges.append((first(16 * b1 + q), first(16 * b1 + q + 1)))
    for b2 in range(nblob // 32):
        lo, hi = 2 * b2, 2 * b2 + 1
        for c in range(8):
            edges.append((first(16 * lo + 2 * c) + 3, first(16 * hi + 2 * c) + 3))
            if c + 1 < 8:
                edges.append((first(16 * hi + 2 * c) + 5, first(16 * lo + 2 * c + 2) + 5))
    e = np.sort(np.array(edges, np.int64), axis=1)
    ne = e.shape[0]
    w = rng.uniform(0.5, 2.0, size=ne)
    off = -w[:, None, None] * np.eye(3) + 0.1 * rng.normal(size=(ne, 3, 3))
    rowsum = np.zeros(nv)
    absn = np.abs(off).sum(axis=2).max(axis=1)
    np.add.at(rowsum, e[:, 0], absn)
    np.add.at(rowsum, e[:, 1], absn)
    S = rng.normal(size=(nv, 3, 3)) * 0.1
    diag = (rowsum + 1.0)[:, None, None] * np.eye(3) + 0.5 * (S + S.transpose(0, 2, 1))
    rows = np.concatenate([np.arange(nv), e[:, 0]])
    cols = np.concatenate([np.arange(nv), e[:, 1]])
    vals = np.concatenate([diag, off]).reshape(-1, 9)
    A = O.bsr_from_upper_triplets(nv, rows, cols, vals)
    part = (np.arange(nv) // 16).astype(np.int32)
    return A, part, rng.normal(size=3 * nv)


def bring_into_contact(scene, a, b, method="bvh", gap_frac=0.5):
    """Positions with body b translated towards body a along the line through
    their centroids until the bodies are within dhat of each other: bisection
    (on the oracle's contact pair count, which must exceed its value before
    the motion) for the first translation t* with a new pair closer than dhat,
    then gap_frac * dhat further, so the separation lies
    in [(1 - gap_frac) dhat, dhat) (a translation by s moves every point by s)."""
    import dataclasses
    import pyoracle as O
    x = scene.x.copy()
    ia, ib = scene.body_id == a, scene.body_id == b
    # the contact search sees only the primitives of the two bodies
    keep = ia | ib
    sub = dataclasses.replace(
        scene, faces=scene.faces[keep[scene.faces].all(1)],
        edges=scene.edges[keep[scene.edges].all(1)],
        surf_verts=scene.surf_verts[keep[scene.surf_verts]])
    d = x[ia].mean(0) - x[ib].mean(0)
    length = float(np.linalg.norm(d))
    u = d / length

    def npairs(t):
        y = x.copy()
        y[ib] += t * u
        _, c = O.contact_pairs(sub, y, method)
        return int(c[0])

    base = npairs(0.0)  # pairs that do not involve the motion of b
    lo, hi = 0.0, length
    assert npairs(hi) > base
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if npairs(mid) > base:
            hi = mid
        else:
            lo = mid
    x[ib] += (hi + gap_frac * scene.dhat) * u
    return x


def lower_onto_ground(x, dhat, frac=0.5):
    """Translate every vertex down so the lowest sits frac * dhat above y = 0."""
    y = x.copy()
    y[:, 1] -= y[:, 1].min() - frac * dhat
    return y
