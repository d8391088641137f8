"""Pins of the oracle's distance classification, pair enumeration and the
GIPC barrier (closed forms, sympy, brute force, finite differences)."""
import numpy as np
import pytest


def _seg_pt(p, a, b):
    t = np.clip((p - a) @ (b - a) / ((b - a) @ (b - a)), 0, 1)
    return a + t * (b - a), t


def _closest_on_triangle(p, a, b, c, n=400):
    """brute force: dense barycentric grid + edge/vertex candidates"""
    best, where = None, None
    for u in np.linspace(0, 1, n):
        v = np.linspace(0, 1 - u, max(2, int(n * (1 - u))))
        q = a[None] + u * (b - a)[None] + v[:, None] * (c - a)[None]
        d = ((q - p) ** 2).sum(1)
        k = d.argmin()
        if best is None or d[k] < best:
            best = d[k]
    return best


def test_distance_functions_vs_numpy(oracle):
    L = oracle.lib()
    P = oracle.P
    rng = np.random.default_rng(1)
    for _ in range(50):
        v = rng.normal(size=(4, 3))
        n = np.cross(v[2] - v[1], v[3] - v[1])
        n /= np.linalg.norm(n)
        assert L.orc_d_pt(*[P(np.ascontiguousarray(a)) for a in v]) == pytest.approx(
            ((v[0] - v[1]) @ n) ** 2, rel=1e-10)
        e = v[2] - v[1]
        t = (v[0] - v[1]) @ e / (e @ e)
        assert L.orc_d_pe(*[P(np.ascontiguousarray(a)) for a in v[:3]]) == pytest.approx(
            ((v[0] - v[1] - t * e) ** 2).sum(), rel=1e-10)
        # line-line distance by least squares
        A = np.stack([v[1] - v[0], -(v[3] - v[2])], 1)
        st = np.linalg.lstsq(A, v[2] - v[0], rcond=None)[0]
        d = v[0] + st[0] * (v[1] - v[0]) - v[2] - st[1] * (v[3] - v[2])
        assert L.orc_d_ee(*[P(np.ascontiguousarray(a)) for a in v]) == pytest.approx(d @ d, rel=1e-9)


def test_point_triangle_classification_vs_bruteforce(oracle):
    L = oracle.lib()
    P = oracle.P
    rng = np.random.default_rng(2)
    dist = {0: lambda v: ((v[0] - v[1]) ** 2).sum(), 1: lambda v: ((v[0] - v[2]) ** 2).sum(),
            2: lambda v: ((v[0] - v[3]) ** 2).sum()}
    checked = 0
    for _ in range(60):
        v = rng.normal(size=(4, 3))
        t = L.orc_dtype_pt(*[P(np.ascontiguousarray(a)) for a in v])
        true = _closest_on_triangle(v[0], v[1], v[2], v[3])
        if t in dist:
            got = dist[t](v)
        elif t == 6:
            got = L.orc_d_pt(*[P(np.ascontiguousarray(a)) for a in v])
        else:
            ab = {3: (1, 2), 4: (2, 3), 5: (3, 1)}[t]
            got = L.orc_d_pe(P(np.ascontiguousarray(v[0])), P(np.ascontiguousarray(v[ab[0]])),
                             P(np.ascontiguousarray(v[ab[1]])))
        # the classified distance is the true point-triangle distance
        assert got == pytest.approx(true, rel=2e-3, abs=1e-6)
        checked += 1
    assert checked == 60


def test_edge_edge_classification_vs_bruteforce(oracle):
    L = oracle.lib()
    P = oracle.P
    rng = np.random.default_rng(4)
    for _ in range(60):
        v = rng.normal(size=(4, 3))
        t = L.orc_dtype_ee(*[P(np.ascontiguousarray(a)) for a in v])
        # segment-segment distance by a coarse grid refined around its minimum
        lo_s, hi_s, lo_t, hi_t = 0.0, 1.0, 0.0, 1.0
        for _ in range(6):
            s = np.linspace(lo_s, hi_s, 201)
            tt = np.linspace(lo_t, hi_t, 201)
            pa = v[0][None] + s[:, None] * (v[1] - v[0])[None]
            pb = v[2][None] + tt[:, None] * (v[3] - v[2])[None]
            D = ((pa[:, None, :] - pb[None, :, :]) ** 2).sum(-1)
            i, j = np.unravel_index(D.argmin(), D.shape)
            true = D[i, j]
            hs, ht = (hi_s - lo_s) / 200 * 3, (hi_t - lo_t) / 200 * 3
            lo_s, hi_s = max(0.0, s[i] - hs), min(1.0, s[i] + hs)
            lo_t, hi_t = max(0.0, tt[j] - ht), min(1.0, tt[j] + ht)
        pp = lambda i, j: ((v[i] - v[j]) ** 2).sum()
        pe = lambda i, j, k: L.orc_d_pe(P(np.ascontiguousarray(v[i])), P(np.ascontiguousarray(v[j])),
                                        P(np.ascontiguousarray(v[k])))
        got = {0: lambda: pp(0, 2), 1: lambda: pp(0, 3), 2: lambda: pe(0, 2, 3),
               3: lambda: pp(1, 2), 4: lambda: pp(1, 3), 5: lambda: pe(1, 2, 3),
               6: lambda: pe(2, 0, 1), 7: lambda: pe(3, 0, 1),
               8: lambda: L.orc_d_ee(*[P(np.ascontiguousarray(a)) for a in v])}[t]()
        assert got == pytest.approx(true, rel=1e-7, abs=1e-12)


def test_barrier_lambda0_sympy(oracle):
    sp = pytest.importorskip("sympy")
    s, k, dh = sp.symbols("s k dh", positive=True)
    I = s ** 2
    b = k * dh ** 2 * (I - 1) ** 2 * sp.log(I) ** 2
    d2b = sp.diff(b, s, 2)
    for sv in (0.05, 0.3, 0.7, 0.95):
        ref = float(d2b.subs({s: sv, k: 2.5, dh: 1e-6}))
        got = oracle.lib().orc_barrier_lambda0(sv * sv, 1e-6, 2.5)
        assert got == pytest.approx(ref, rel=1e-9)


def _codes():
    # (positions, code) with the contact primitive at ~0.4 dhat
    x = np.array([[0.0, 4e-4, 0.0], [-0.1, 0.0, -0.1], [0.1, 0.0, -0.1], [0.0, 0.0, 0.15]])
    yield "PT", x, np.array([-1, 1, 2, 3], np.int32)
    yield "PE", x, np.array([-1, 1, 2, -1], np.int32)
    x2 = np.array([[0.0, 4e-4, 0.0], [0.0, 0.0, 0.0], [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])
    yield "PP", x2, np.array([-1, 1, -1, -1], np.int32)
    x3 = np.array([[-0.1, 5e-4, 0.01], [0.1, 5e-4, -0.02], [0.02, 0.0, -0.1], [-0.03, 0.0, 0.1]])
    yield "EE", x3, np.array([0, 1, 2, 3], np.int32)


@pytest.mark.parametrize("case", list(range(4)))
def test_barrier_gradient_hessian(oracle, case):
    name, x, code = list(_codes())[case]
    dhat2, kappa = 1e-6, 7.0
    ids, g, H = oracle.barrier_grad_hess(x.ravel(), code, dhat2, kappa)
    nv = ids.size
    eps = 1e-9
    gfd = np.zeros(3 * nv)
    sfd = np.zeros(3 * nv)
    L = oracle.lib()
    P = oracle.P

    def s_of(xx):
        v = [np.ascontiguousarray(xx[i]) for i in ids]
        d2 = {"PT": lambda: L.orc_d_pt(*map(P, v)), "PE": lambda: L.orc_d_pe(*map(P, v)),
              "PP": lambda: L.orc_d_pp(*map(P, v)), "EE": lambda: L.orc_d_ee(*map(P, v))}[name]()
        return np.sqrt(d2 / dhat2)

    for a in range(nv):
        for c in range(3):
            xp, xm = x.copy(), x.copy()
            xp[ids[a], c] += eps
            xm[ids[a], c] -= eps
            gfd[3 * a + c] = (oracle.barrier_energy(xp.ravel(), code, dhat2, kappa)
                              - oracle.barrier_energy(xm.ravel(), code, dhat2, kappa)) / (2 * eps)
            sfd[3 * a + c] = (s_of(xp) - s_of(xm)) / (2 * eps)
    assert np.allclose(g, gfd, rtol=1e-5, atol=1e-7 * np.abs(g).max())
    sv = s_of(x)
    lam0 = oracle.lib().orc_barrier_lambda0(sv * sv, dhat2, kappa)
    assert np.allclose(H, lam0 * np.outer(sfd, sfd), rtol=1e-5, atol=1e-7 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()


@pytest.mark.parametrize("seed", [0, 1])
def test_pairs_bvh_equals_bruteforce(oracle, seed):
    from sgipc import scene as S
    sc = S.stack_scene(n=3, cells=2, gap=4e-4, rot_deg=11.0 + seed)
    rng = np.random.default_rng(seed)
    x = sc.x + 1e-4 * rng.normal(size=sc.x.shape)
    pb, cb = oracle.contact_pairs(sc, x, "brute")
    pv, cv = oracle.contact_pairs(sc, x, "bvh")
    assert cb[0] > 0
    assert sorted(map(tuple, pb.tolist())) == sorted(map(tuple, pv.tolist()))
    np.testing.assert_array_equal(cb, cv)
    # every record's primitive distance is below dhat^2 (brute-force recheck)
    L = oracle.lib()
    P = oracle.P
    for c in pb:
        if c[0] >= 0:
            d = L.orc_d_ee(*[P(np.ascontiguousarray(x[i])) for i in c])
        else:
            ids = [-c[0] - 1, c[1]] + [i for i in c[2:] if i >= 0]
            if c[2] < 0:
                ids = ids[:2]
            elif c[3] < 0:
                ids = ids[:3]
            f = {2: L.orc_d_pp, 3: L.orc_d_pe, 4: L.orc_d_pt}[len(ids)]
            d = f(*[P(np.ascontiguousarray(x[i])) for i in ids])
        assert d < sc.dhat ** 2


def test_accd_bounds(oracle):
    # head-on point vs triangle: true time of impact known in closed form
    p = np.array([0.0, 1.0, 0.0])
    t0, t1, t2 = np.array([-1.0, 0, -1]), np.array([1.0, 0, -1]), np.array([0.0, 0, 1])
    dp = np.array([0.0, -2.0, 0.0])
    z = np.zeros(3)
    toc = oracle.pt_ccd(p, t0, t1, t2, dp, z, z, z, eta=0.2)
    assert 0.0 < toc <= 0.5 and toc >= 0.5 * 0.8 * 0.99  # stops between (1-eta) and 1 of TOI
    # separating motion never collides
    assert oracle.pt_ccd(p, t0, t1, t2, -dp, z, z, z) == 1.0
    # crossing edges
    a0, a1 = np.array([-1.0, 1.0, 0]), np.array([1.0, 1.0, 0])
    b0, b1 = np.array([0.0, 0.0, -1]), np.array([0.0, 0.0, 1])
    da = np.array([0.0, -4.0, 0])
    toc = oracle.ee_ccd(a0, a1, b0, b1, da, da, z, z, eta=0.2)
    assert 0.0 < toc <= 0.25 and toc >= 0.25 * 0.8 * 0.99
