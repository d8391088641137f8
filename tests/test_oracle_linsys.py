"""Pins of the oracle's BSR merge, MAS hierarchy/operator and PCG against
independent numpy/BFS computations (no GPU)."""
import numpy as np
import pytest

from helpers import assembled_system


def test_bsr_merge_vs_dense(oracle):
    rng = np.random.default_rng(0)
    nb, nt = 20, 300
    r = rng.integers(0, nb, nt)
    c = rng.integers(0, nb, nt)
    r, c = np.minimum(r, c), np.maximum(r, c)
    v = rng.normal(size=(nt, 3, 3))
    A = oracle.bsr_from_upper_triplets(nb, r, c, v)
    D = np.zeros((3 * nb, 3 * nb))
    for k in range(nt):
        D[3 * r[k]:3 * r[k] + 3, 3 * c[k]:3 * c[k] + 3] += v[k]
        if r[k] != c[k]:
            D[3 * c[k]:3 * c[k] + 3, 3 * r[k]:3 * r[k] + 3] += v[k].T
    assert np.allclose(A.dense(), D, rtol=0, atol=1e-14)
    x = rng.normal(size=3 * nb)
    assert np.allclose(A.spmv(x), D @ x, rtol=1e-14, atol=1e-13)


def _hierarchy_bfs(A, part_id, levels):
    """independent: components per 16-node bank by BFS, numbered bank by
    bank in order of each component's lowest lane"""
    nv = A.nb
    edges = [(i, j) for i in range(nv) for k in range(A.rowptr[i], A.rowptr[i + 1])
             for j in [A.cols[k]] if j != i]
    slot = np.zeros(nv, int)
    prev, idx = None, 0
    for v in range(nv):
        if part_id[v] != prev:
            prev, idx = part_id[v], 0
        slot[v] = 16 * part_id[v] + idx
        idx += 1
    node = slot.copy()
    nnodes = 16 * (part_id.max() + 1)
    valid = np.zeros(nnodes, bool)
    valid[slot] = True
    chains = []
    for lvl in range(1, levels):
        adj = {}
        for i, j in edges:
            a, b = node[i], node[j]
            if a != b and a // 16 == b // 16:
                adj.setdefault(a, set()).add(b)
                adj.setdefault(b, set()).add(a)
        newid = -np.ones(nnodes, int)
        cnt = 0
        for a in range(nnodes):
            if not valid[a] or newid[a] >= 0:
                continue
            stack, comp = [a], {a}
            while stack:
                u = stack.pop()
                for w in adj.get(u, ()):
                    if w not in comp:
                        comp.add(w)
                        stack.append(w)
            for u in comp:
                newid[u] = cnt
            cnt += 1
        node = newid[node]
        chains.append(node.copy())
        nnodes = cnt
        valid = np.ones(cnt, bool)
    return slot, chains


@pytest.mark.parametrize("scene_name", ["bar", "stack"])
def test_mas_hierarchy_is_bank_components(oracle, scene_name):
    from sgipc import scene as S
    sc = S.bar_scene() if scene_name == "bar" else S.stack_scene(n=3, cells=3)
    w, g, A = assembled_system(sc)
    m = oracle.Mas(A, sc.part_id, fp32=False)
    slot, chains = _hierarchy_bfs(A, sc.part_id, m.levels)
    sizes = m.level_sizes()
    ch = m.chain()
    for l in range(1, m.levels):
        np.testing.assert_array_equal(ch[:, l - 1] - sizes[l, 1], chains[l - 1])


def test_mas_operator_dense(oracle):
    """z = sum_l P_l (P_l^T A P_l restricted to banks)^-1 P_l^T r, built densely"""
    from sgipc import scene as S
    sc = S.stack_scene(n=2, cells=2)
    w, g, A = assembled_system(sc, perturb=1e-3)
    m = oracle.Mas(A, sc.part_id, fp32=False)
    D = A.dense()
    nv = A.nb
    ch = m.chain()
    sizes = m.level_sizes()
    slot, _ = _hierarchy_bfs(A, sc.part_id, m.levels)
    M = np.zeros_like(D)
    ids = [slot] + [ch[:, l - 1] for l in range(1, m.levels)]
    for l, gid in enumerate(ids):
        banks = gid // 16
        for b in np.unique(banks):
            members = np.where(banks == b)[0]
            nodes = np.unique(gid[members])
            Pm = np.zeros((3 * nv, 3 * len(nodes)))
            for k, nd in enumerate(nodes):
                for v in members[gid[members] == nd]:
                    Pm[3 * v:3 * v + 3, 3 * k:3 * k + 3] = np.eye(3)
            B = Pm.T @ D @ Pm
            M += Pm @ np.linalg.inv(B) @ Pm.T
    rng = np.random.default_rng(2)
    for _ in range(3):
        r = rng.normal(size=3 * nv)
        assert np.allclose(m.apply(r), M @ r, rtol=1e-9, atol=1e-9 * np.abs(M @ r).max())
    assert np.abs(M - M.T).max() <= 1e-9 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_pcg_vs_dense_solve(oracle):
    from sgipc import scene as S
    sc = S.bar_scene()
    w, g, A = assembled_system(sc, perturb=2e-3)
    D = A.dense()
    m = oracle.Mas(A, sc.part_id, fp32=False)
    x, it, rz0, rz = oracle.pcg(A, g, m, 1, rel_tol=1e-10)
    ref = np.linalg.solve(D, g)
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-7
    assert it % 10 == 0  # the stop test runs every 10 iterations
    x1, it1, _, _ = oracle.pcg(A, g, m, 1, rel_tol=1e-4)
    x2, it2, _, _ = oracle.pcg(A, g, None, 0, rel_tol=1e-4)
    assert it1 <= it2  # MAS never needs more iterations than block Jacobi here
