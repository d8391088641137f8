"""Pins of the oracle's Morton codes and LBVH: bit interleave from first
principles, Karras invariants checked by brute force (no GPU)."""
import numpy as np
import pytest


def _interleave(v):
    """bit i of v -> bit 3 i (textbook 10-bit dilation)"""
    r = 0
    for i in range(10):
        r |= ((v >> i) & 1) << (3 * i)
    return r


def test_expand_bits_all_10bit(oracle):
    L = oracle.lib()
    for v in range(1024):
        assert L.orc_expand_bits(v) == _interleave(v)


def test_morton_code_clamp_and_order(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rng.uniform(-0.2, 1.2, 3)
        q = np.clip(np.floor(np.clip(p * 1024.0, 0.0, 1023.0)), 0, 1023).astype(int)
        expect = (_interleave(q[0]) << 2) + (_interleave(q[1]) << 1) + _interleave(q[2])
        assert L.orc_morton_code(*p) == expect


def _check_tree(keys, nodes, aabbs, leaf_boxes):
    n = keys.size
    low = (keys & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert np.all(np.diff(keys.astype(object)) > 0) if n > 1 else True
    assert sorted(low.tolist()) == list(range(n))
    NONE = 0xFFFFFFFF
    # leaves
    for i in range(n):
        nd = nodes[n - 1 + i]
        assert nd[1] == NONE and nd[2] == NONE and nd[3] == low[i]
        np.testing.assert_array_equal(aabbs[n - 1 + i], leaf_boxes[low[i]])
    # ranges by recursion from the root; children partition the parent range
    rng_of = {}

    def walk(u):
        if u >= n - 1:
            rng_of[u] = (u - (n - 1), u - (n - 1))
            return rng_of[u]
        L, R = int(nodes[u][1]), int(nodes[u][2])
        assert nodes[L][0] == u and nodes[R][0] == u
        a, b = walk(L), walk(R)
        assert a[1] + 1 == b[0]
        rng_of[u] = (a[0], b[1])
        # Karras split: highest differing bit between the range's end keys
        f, l = a[0], b[1]
        kf, kl = int(keys[f]), int(keys[l])
        top = (kf ^ kl).bit_length() - 1
        assert (int(keys[a[1]]) >> top) & 1 == (kf >> top) & 1
        assert (int(keys[b[0]]) >> top) & 1 == (kl >> top) & 1
        # refit: union of the children
        np.testing.assert_array_equal(aabbs[u][:3], np.maximum(aabbs[L][:3], aabbs[R][:3]))
        np.testing.assert_array_equal(aabbs[u][3:], np.minimum(aabbs[L][3:], aabbs[R][3:]))
        return rng_of[u]

    if n > 1:
        assert nodes[0][0] == NONE
        assert walk(0) == (0, n - 1)
        assert len(rng_of) == 2 * n - 1


@pytest.mark.parametrize("which", ["faces", "edges"])
def test_lbvh_invariants_bar(oracle, which):
    from sgipc import scene as S
    sc = S.bar_scene()
    prims = sc.faces if which == "faces" else sc.edges
    keys, nodes, aabbs = oracle.lbvh(sc.x, prims)
    lb = np.zeros((prims.shape[0], 6))
    for e, pr in enumerate(prims):
        p = sc.x[pr]
        lb[e, :3] = p.max(0)
        lb[e, 3:] = p.min(0)
    _check_tree(keys, nodes, aabbs, lb)
    # keys: morton of the leaf-box centre relative to the scene box
    lo, hi = lb[:, 3:].min(0), lb[:, :3].max(0)
    c = (lb[:, :3] + lb[:, 3:]) * 0.5
    L = oracle.lib()
    mc = np.array([L.orc_morton_code(*((c[e] - lo) / (hi - lo))) for e in range(len(lb))],
                  dtype=np.uint64)
    expect = np.sort((mc << np.uint64(32)) | np.arange(len(lb), dtype=np.uint64))
    np.testing.assert_array_equal(keys, expect)


def test_lbvh_invariants_random_cloud(oracle):
    rng = np.random.default_rng(3)
    x = rng.uniform(size=(600, 3))
    prims = rng.integers(0, 600, size=(400, 3)).astype(np.uint32)
    keys, nodes, aabbs = oracle.lbvh(x, prims)
    lb = np.stack([np.concatenate([x[p].max(0), x[p].min(0)]) for p in prims])
    _check_tree(keys, nodes, aabbs, lb)
