"""Pins of the oracle's cloth terms (femEnergy.cu:262-309, 360-412, 565-740,
1003-1041, 2215-2300, 2431-2630; main.cu:303-307; load_mesh.cpp:320-379):
closed forms of the stiffness parameters, the rest frame and the
Baraff-Witkin energy, finite differences of every gradient and unprojected
Hessian, the projected Hessians against numpy eigen-projections, the
dihedral angle against the reference's normal-based formula, the hinge list
against a brute-force enumeration, and physical behaviour of small sheets."""
import numpy as np
import pytest


def _psd(H):
    w, V = np.linalg.eigh(0.5 * (H + H.T))
    return (V * np.maximum(w, 0)) @ V.T


def test_cloth_params_closed_form(oracle):
    E, nu, Eb, h, r = 1e6, 0.3, 2e5, 1e-3, 0.5
    st, sh, bd = oracle.cloth_params(E, nu, Eb, h, r)
    assert st == pytest.approx(E / 2.6, rel=1e-15)
    assert sh == pytest.approx(0.03 * E / 2.6 * r, rel=1e-15)
    assert bd == pytest.approx(Eb * 1e-9 / (24 * (1 - 0.09)), rel=1e-15)


def _rodrigues_angle(n):
    """minimal rotation taking unit n to +z, from axis and angle"""
    z = np.array([0.0, 0.0, 1.0])
    k = np.cross(n, z)
    s = np.linalg.norm(k)
    c = n @ z
    if s < 1e-300:
        return np.eye(3) if c > 0 else np.diag([-1.0, -1.0, 1.0])
    k /= s
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + s * K + (1 - c) * K @ K


@pytest.mark.parametrize("seed", range(5))
def test_tri_rest_frame(oracle, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(3, 3))
    if seed == 3:
        x[:, 2] = 0.2                  # normal +z: the frame is the xy plane itself
    if seed == 4:
        x[:, 2] = 0.2
        x = x[[0, 2, 1]]               # normal -z: R = diag(-1, -1, 1)
    Dm_inv, area = oracle.tri_precompute(x, [0, 1, 2])
    e1, e2 = x[1] - x[0], x[2] - x[0]
    nrm = np.cross(e1, e2)
    assert area == pytest.approx(0.5 * np.linalg.norm(nrm), rel=1e-14)
    R = _rodrigues_angle(nrm / np.linalg.norm(nrm))
    Dm = np.stack([(R @ e1)[:2], (R @ e2)[:2]], 1)
    np.testing.assert_allclose(Dm_inv, np.linalg.inv(Dm), rtol=1e-10, atol=1e-12)
    # at rest F = Ds Dm^-1 has orthonormal columns
    F = np.stack([e1, e2], 1) @ Dm_inv
    np.testing.assert_allclose(F.T @ F, np.eye(2), atol=1e-12)


def test_bw_closed_forms(oracle):
    st, sh, r = 3.0, 0.7, 0.4
    for s in (0.8, 1.0, 1.25, 1.6):
        F = np.array([[s, 0.0], [0.0, 1.0], [0.0, 0.0]])
        cub = r * (s - 1) ** 3 if s > 1 else 0.0
        assert oracle.bw_psi(F, st, sh, r) == pytest.approx(st * ((s - 1) ** 2 + cub), abs=1e-15)
    # pure shear: columns (1,0,0), (g,1,0): I6 = g, |F v| = sqrt(1+g^2)
    g = 0.3
    F = np.array([[1.0, g], [0.0, 1.0], [0.0, 0.0]])
    sv = np.sqrt(1 + g * g)
    want = st * ((sv - 1) ** 2 + r * (sv - 1) ** 3) + sh * g * g
    assert oracle.bw_psi(F, st, sh, r) == pytest.approx(want, rel=1e-14)
    # rest state: zero energy and stress
    F0 = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]])
    assert oracle.bw_psi(F0, st, sh, r) == 0.0
    assert np.abs(oracle.bw_pk1(F0, st, sh, r)).max() == 0.0


def _vecF(F):
    return np.concatenate([F[:, 0], F[:, 1]])


def _unvec(v):
    return np.stack([v[:3], v[3:]], 1)


@pytest.mark.parametrize("seed", range(6))
def test_bw_derivatives_fd(oracle, seed):
    rng = np.random.default_rng(10 + seed)
    F = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]]) + rng.normal(size=(3, 2)) * 0.2
    st, sh, r = 2.0, 0.9, 0.6
    h = 1e-6
    f = _vecF(F)
    Pfd = np.zeros(6)
    for k in range(6):
        e = np.zeros(6)
        e[k] = h
        Pfd[k] = (oracle.bw_psi(_unvec(f + e), st, sh, r) - oracle.bw_psi(_unvec(f - e), st, sh, r)) / (2 * h)
    np.testing.assert_allclose(_vecF(oracle.bw_pk1(F, st, sh, r)), Pfd, rtol=1e-7, atol=1e-8)
    # unprojected Hessian = d(PK1)/dF; projection per term = numpy eigen-projection
    for (a, b) in ((1.0, 0.0), (0.0, 1.0)):
        H = oracle.bw_hess(F, a, b, r, project=False)
        Hfd = np.zeros((6, 6))
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            Hfd[:, k] = (_vecF(oracle.bw_pk1(_unvec(f + e), a, b, r))
                         - _vecF(oracle.bw_pk1(_unvec(f - e), a, b, r))) / (2 * h)
        np.testing.assert_allclose(H, Hfd, rtol=1e-6, atol=1e-7)
        if a:      # the two stretch axes project separately
            Hp = np.zeros((6, 6))
            Hp[:3, :3] = _psd(Hfd[:3, :3])
            Hp[3:, 3:] = _psd(Hfd[3:, 3:])
        else:
            Hp = _psd(Hfd)
        np.testing.assert_allclose(oracle.bw_hess(F, a, b, r, project=True), Hp, rtol=1e-6, atol=1e-6)


def _bw_shear_closed_form(F):
    """the reference's closed-form shear projection (femEnergy.cu:666-740),
    evaluated in numpy: 2 (|I6| (T - Tq Tq^T/|Tq|^2) + lam0 q0 q0^T)"""
    S = np.zeros((6, 6))
    S[:3, 3:] = np.eye(3)
    S[3:, :3] = np.eye(3)
    g = np.concatenate([F[:, 1], F[:, 0]])
    I6 = F[:, 0] @ F[:, 1]
    I2 = (F * F).sum()
    lam0 = 0.5 * (I2 + np.sqrt(I2 * I2 + 12 * I6 * I6))
    q0 = I6 * S @ g + lam0 * g
    q0 /= np.linalg.norm(q0)
    T = 0.5 * (np.eye(6) + (1.0 if I6 >= 0 else -1.0) * S)
    Tq = T @ q0
    return 2 * (abs(I6) * (T - np.outer(Tq, Tq) / (Tq @ Tq)) + lam0 * np.outer(q0, q0))


@pytest.mark.parametrize("seed", range(4))
def test_shear_projection_matches_closed_form(oracle, seed):
    rng = np.random.default_rng(30 + seed)
    F = np.array([[1.0, 0.0], [0.0, 1.0], [0.0, 0.0]]) + rng.normal(size=(3, 2)) * 0.3
    np.testing.assert_allclose(oracle.bw_hess(F, 0.0, 1.0, 0.0), _bw_shear_closed_form(F),
                               rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("seed", range(3))
def test_tri_grad_hess_fd(oracle, seed):
    rng = np.random.default_rng(40 + seed)
    xr = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.01], [0.02, 0.0, 0.09]])
    Dm, area = oracle.tri_precompute(xr, [0, 1, 2])
    st, sh, r, dt = 5e3, 30.0, 0.5, 0.01
    x = xr + rng.normal(size=(3, 3)) * 0.02
    g, H = oracle.tri_grad_hess(x, [0, 1, 2], Dm, area, st, sh, r, dt)
    E = lambda y: dt * dt * oracle.tri_energy(y, [0, 1, 2], Dm, area, st, sh, r)
    h = 1e-7
    gfd = np.zeros(9)
    for k in range(9):
        e = np.zeros(9)
        e[k] = h
        gfd[k] = (E((x.ravel() + e).reshape(3, 3)) - E((x.ravel() - e).reshape(3, 3))) / (2 * h)
    np.testing.assert_allclose(g, gfd, rtol=1e-6, atol=1e-9 * np.abs(g).max())
    np.testing.assert_allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() > -1e-9 * np.abs(H).max()
    # translation invariance: the three vertex gradients sum to zero
    np.testing.assert_allclose(g.reshape(3, 3).sum(0), 0, atol=1e-12 * np.abs(g).max())


def test_tri_hess_unprojected_regime_fd(oracle):
    """both columns stretched and orthogonal (I6 = 0): no term is projected,
    so the Hessian is the finite difference of the gradient"""
    xr = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [0.0, 0.1, 0.0]])
    Dm, area = oracle.tri_precompute(xr, [0, 1, 2])
    x = xr * np.array([1.2, 1.1, 1.0])
    st, sh, r, dt = 5e3, 30.0, 0.5, 0.01
    g, H = oracle.tri_grad_hess(x, [0, 1, 2], Dm, area, st, sh, r, dt)
    h = 1e-7
    Hfd = np.zeros((9, 9))
    for k in range(9):
        e = np.zeros(9)
        e[k] = h
        gp = oracle.tri_grad_hess((x.ravel() + e).reshape(3, 3), [0, 1, 2], Dm, area, st, sh, r, dt)[0]
        gm = oracle.tri_grad_hess((x.ravel() - e).reshape(3, 3), [0, 1, 2], Dm, area, st, sh, r, dt)[0]
        Hfd[:, k] = (gp - gm) / (2 * h)
    # the shear term at I6 = 0 is 2 g g^T (PSD), so FD matches exactly there
    np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-6 * np.abs(H).max())


def _theta_ref(q0, q1, q2, q3):
    """LibShell::edgeTheta / angle (femEnergy.cu:102-170) in numpy"""
    n0 = np.cross(q0 - q2, q1 - q2)
    n1 = np.cross(q1 - q3, q0 - q3)
    a = q1 - q0
    return 2 * np.arctan2(np.cross(n0, n1) @ a / np.linalg.norm(a),
                          n0 @ n1 + np.linalg.norm(n0) * np.linalg.norm(n1))


def _hinge(fold, rng=None):
    q0, q1 = np.array([0.0, 0.0, 0.0]), np.array([0.0, 0.0, 0.1])
    q2 = np.array([0.08, 0.0, 0.04])
    q3 = np.array([-0.08 * np.cos(fold), 0.08 * np.sin(fold), 0.06])
    q = np.stack([q0, q1, q2, q3])
    if rng is not None:
        q = q + rng.normal(size=q.shape) * 0.005
    return q


@pytest.mark.parametrize("fold", [0.0, 0.3, -0.7, 1.2, 2.5])
def test_dihedral_matches_reference_formula(oracle, fold):
    q = _hinge(fold)
    t = oracle.dihedral(*q, derivs=False)
    assert t == pytest.approx(_theta_ref(*q), abs=1e-14)
    # planar-angle closed form: |theta| = fold for this hinge
    assert abs(t) == pytest.approx(abs(fold), abs=1e-12)


@pytest.mark.parametrize("seed", range(5))
def test_dihedral_derivatives_fd(oracle, seed):
    rng = np.random.default_rng(50 + seed)
    q = _hinge([0.2, -0.5, 1.0, 2.0, -2.6][seed], rng)
    t, g, H = oracle.dihedral(*q)
    h = 1e-7
    gfd = np.zeros(12)
    Hfd = np.zeros((12, 12))
    for k in range(12):
        e = np.zeros(12)
        e[k] = h
        qp, qm = (q.ravel() + e).reshape(4, 3), (q.ravel() - e).reshape(4, 3)
        gfd[k] = (oracle.dihedral(*qp, derivs=False) - oracle.dihedral(*qm, derivs=False)) / (2 * h)
        Hfd[:, k] = (oracle.dihedral(*qp)[1] - oracle.dihedral(*qm)[1]) / (2 * h)
    np.testing.assert_allclose(g, gfd, rtol=1e-6, atol=1e-7 * np.abs(g).max())
    np.testing.assert_allclose(H, Hfd, rtol=1e-5, atol=1e-5 * np.abs(H).max())
    np.testing.assert_allclose(H, H.T, atol=1e-10 * np.abs(H).max())
    # rigid invariance: translations and rotations leave theta unchanged
    G = g.reshape(4, 3)
    np.testing.assert_allclose(G.sum(0), 0, atol=1e-10 * np.abs(g).max())
    np.testing.assert_allclose(np.cross(q, G).sum(0), 0, atol=1e-10 * np.abs(g).max())


@pytest.mark.parametrize("seed", range(3))
def test_bend_grad_hess(oracle, seed):
    rng = np.random.default_rng(60 + seed)
    q0 = _hinge(0.4)
    theta0 = oracle.dihedral(*q0, derivs=False)
    len0 = 0.1
    k, dt = 0.05, 0.01
    q = _hinge(0.4 + [0.3, -0.5, 0.9][seed], rng)
    hinge = np.array([0, 1, 2, 3], np.uint32)
    g, H = oracle.bend_grad_hess(q, hinge, theta0, len0, k, dt)
    E = lambda y: dt * dt * oracle.bend_energy(y, hinge, theta0, len0, k)
    h = 1e-7
    gfd = np.zeros(12)
    for j in range(12):
        e = np.zeros(12)
        e[j] = h
        gfd[j] = (E((q.ravel() + e).reshape(4, 3)) - E((q.ravel() - e).reshape(4, 3))) / (2 * h)
    np.testing.assert_allclose(g, gfd, rtol=1e-6, atol=1e-8 * np.abs(g).max())
    t, gt, Ht = oracle.dihedral(*q)
    s = k * len0 * dt * dt
    Hraw = 2 * ((t - theta0) * Ht + np.outer(gt, gt))
    np.testing.assert_allclose(H, s * _psd(Hraw), rtol=1e-8, atol=1e-9 * np.abs(H).max())
    assert E(q) == pytest.approx(dt * dt * k * len0 * (t - theta0) ** 2, rel=1e-14)


def test_bend_hinges_brute_force(oracle):
    from sgipc import scene as S
    _, tris, _ = S.cloth_grid(6, 1.0)
    got = oracle.bend_hinges(tris)
    opp = {}
    for f, t in enumerate(tris.tolist()):
        for k in range(3):
            a, b, o = t[k], t[(k + 1) % 3], t[(k + 2) % 3]
            opp.setdefault((min(a, b), max(a, b)), []).append(o)
    want = [(a, b, v[0], v[1]) for (a, b), v in sorted(opp.items()) if len(v) == 2]
    assert [tuple(r) for r in got.tolist()] == want
    # 2 (n-1)^2 triangles of an n x n grid have 3 (n-1)^2 - 2 (n-1) interior edges
    assert len(want) == 3 * 25 - 10


def _sheet(n=5, size=0.2, **kw):
    from sgipc import scene as S
    x, tris, part = S.cloth_grid(n, size, origin=(0.0, 0.5, 0.0))
    defaults = dict(dt=0.01, dhat=1e-3, cloth_E=1e5, cloth_nu=0.3, cloth_bend_E=1e5,
                    cloth_thickness=1e-3, cloth_density=200.0, cloth_strain_rate=0.1,
                    use_ground=False)
    defaults.update(kw)
    bt = kw.pop("btype", 0)
    sc = S.assemble_bodies([], cloths=[(x, tris, part, bt)], **{k: v for k, v in defaults.items()
                                                                 if k != "btype"})
    sc.abs_xdelta_tol = 1e-6
    return sc


def test_sheet_rest_state_is_stress_free(oracle):
    """a flat sheet in free fall stays undeformed: every element gradient
    vanishes at rest, so x_n = x_0 + g dt^2 (1 + 2 + ... + n) exactly"""
    sc = _sheet()
    w = oracle.World(sc)
    grad = w.assemble()[0]
    # at rest, only the kinetic term m (x - x~) = -m g dt^2 remains
    mass_g = grad.reshape(-1, 3)
    assert np.abs(mass_g[:, [0, 2]]).max() < 1e-15
    assert np.ptp(mass_g[:, 1] / mass_g[:, 1].mean()) > 0  # lumped masses differ at the rim
    x0 = w.positions().copy()
    for n in range(1, 4):
        w.time_step()
        drop = -9.8 * 0.01 ** 2 * n * (n + 1) / 2
        np.testing.assert_allclose(w.positions(), x0 + np.array([0.0, drop, 0.0]), atol=1e-10)


def test_lumped_mass_and_energy_terms(oracle):
    sc = _sheet(n=4, size=0.3)
    w = oracle.World(sc)
    t = w.energy()
    assert t[6] == pytest.approx(0.0, abs=1e-20)
    # the kinetic term at x = x_prev measures the total mass: 1/2 M |g dt^2|^2
    M = 0.3 * 0.3 * 1e-3 * 200.0
    assert t[1] == pytest.approx(0.5 * M * (9.8e-4) ** 2, rel=1e-12)


def test_cantilever_bending_stiffness(oracle):
    """a strip clamped along two vertex rows bends under gravity: once the
    hinges are stiff enough for a quasi-static response the tip deflection
    scales as 1 / bending stiffness (linear beam theory), and a soft strip
    sags far more"""
    tips = {}
    for Eb in (1e4, 1e8, 1e9):
        sc = _sheet(n=5, size=0.1, cloth_bend_E=Eb, cloth_E=1e6, cloth_thickness=5e-3)
        xr = sc.x.copy()
        clamp = xr[:, 0] < 0.03
        sc.btype = np.where(clamp, 1, 0).astype(np.int32)
        w = oracle.World(sc)
        for _ in range(20):
            w.time_step()
        x = w.positions()
        tips[Eb] = 0.5 - x[np.abs(xr[:, 0] - 0.1) < 1e-9, 1].mean()
        assert np.all(x[clamp] == xr[clamp])
    assert tips[1e8] / tips[1e9] == pytest.approx(10.0, rel=0.2)
    assert tips[1e4] > 10 * tips[1e8] > 0
