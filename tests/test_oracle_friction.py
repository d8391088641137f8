"""Pins of the oracle's lagged friction (GIPC.cu:1254-1804, 5049-5273,
8304-8441; FrictionUtils.cuh): the lagged normal force equals the barrier's
force magnitude (finite differences of the barrier energy in squared
distance), the closest-point coordinates and tangent bases satisfy their
defining orthogonality conditions, gradients and Hessians are the finite
differences of the energy (both the sliding and the C1-mollified static
regime), and a block sliding on the ground decelerates at mu g."""
import numpy as np
import pytest


def _geom(kind, rng, gap):
    """four points in contact configuration of the given kind with distance ~gap"""
    if kind == "PP":
        p = rng.normal(size=3) * 0.1
        return np.stack([p, p + gap * np.array([0.3, 0.9, 0.3]) / np.linalg.norm([0.3, 0.9, 0.3]),
                         p, p]), np.array([-1, 1, -1, -1])
    if kind == "PE":
        a, b = np.array([-0.5, 0.0, 0.1]), np.array([0.6, 0.05, -0.1])
        p = 0.45 * a + 0.55 * b + gap * np.array([0.0, 0.8, 0.6])
        return np.stack([p, a, b, b]), np.array([-1, 1, 2, -1])
    if kind == "PT":
        t0, t1, t2 = np.array([0.0, 0.0, 0.0]), np.array([1.0, 0.1, 0.0]), np.array([0.2, 0.0, 1.0])
        nrm = np.cross(t1 - t0, t2 - t0)
        nrm /= np.linalg.norm(nrm)
        p = 0.3 * t0 + 0.3 * t1 + 0.4 * t2 + gap * nrm
        return np.stack([p, t0, t1, t2]), np.array([-1, 1, 2, 3])
    a0, a1 = np.array([-0.5, 0.0, 0.0]), np.array([0.5, 0.0, 0.02])
    b0, b1 = np.array([0.1, gap, -0.5]), np.array([-0.05, gap, 0.5])
    return np.stack([a0, a1, b0, b1]), np.array([0, 1, 2, 3])


KINDS = ["PP", "PE", "PT", "EE"]


def test_lagged_normal_force_is_barrier_force(oracle):
    """lam = -2 sqrt(s) dE/ds with E(s) = kappa (s - dhat)^2 ln^2(s / dhat)
    (GIPC.cu:703-959 barrier on squared distance s), by central differences."""
    rng = np.random.default_rng(0)
    dhat2, kappa = 1e-6, 3e4
    for kind in KINDS:
        x, code = _geom(kind, rng, 4e-4)
        k, lam, _, _ = oracle.friction_lag(x, code, dhat2, kappa)
        assert k == KINDS.index(kind)
        # squared distance of the pair through the oracle's classified distance
        ids = [0, 1, 2, 3]
        P = [oracle.P(np.ascontiguousarray(x[i])) for i in ids]
        d2 = {"PP": lambda: oracle.lib().orc_d_pp(P[0], P[1]),
              "PE": lambda: oracle.lib().orc_d_pe(P[0], P[1], P[2]),
              "PT": lambda: oracle.lib().orc_d_pt(P[0], P[1], P[2], P[3]),
              "EE": lambda: oracle.lib().orc_d_ee(P[0], P[1], P[2], P[3])}[kind]()
        E = lambda q: kappa * (q - dhat2) ** 2 * np.log(q / dhat2) ** 2
        h = d2 * 1e-6
        dE = (E(d2 + h) - E(d2 - h)) / (2 * h)
        assert lam > 0
        assert lam == pytest.approx(-2 * np.sqrt(d2) * dE, rel=1e-6)


def test_ground_lagged_force_is_barrier_force(oracle):
    """ground: E(s) = -kappa (s - dhat)^2 ln(s / dhat) (GIPC.cu:6839-6860)"""
    dhat2, kappa = 1e-6, 3e4
    n = np.array([0.0, 1.0, 0.0])
    x = np.array([0.1, 5e-4, -0.2])
    lam = oracle.friction_lag_ground(x, n, 0.0, dhat2, kappa)
    s = x[1] ** 2
    E = lambda q: -kappa * (q - dhat2) ** 2 * np.log(q / dhat2)
    h = s * 1e-6
    assert lam == pytest.approx(-2 * np.sqrt(s) * (E(s + h) - E(s - h)) / (2 * h), rel=1e-6)


@pytest.mark.parametrize("kind", KINDS)
def test_basis_and_closest_point(oracle, kind):
    rng = np.random.default_rng(1)
    x, code = _geom(kind, rng, 3e-4)
    x = x + rng.normal(size=x.shape) * 0.01
    _, _, c, B = oracle.friction_lag(x, code, 1e-6, 1.0)
    np.testing.assert_allclose(B.T @ B, np.eye(2), atol=1e-12)
    if kind == "PT":
        e1, e2 = x[2] - x[1], x[3] - x[1]
        r = x[0] - (x[1] + c[0] * e1 + c[1] * e2)
        assert abs(r @ e1) < 1e-12 and abs(r @ e2) < 1e-12
        assert np.all(np.abs(B.T @ np.cross(e1, e2)) < 1e-12)
    elif kind == "EE":
        e01, e23 = x[1] - x[0], x[3] - x[2]
        r = (x[0] + c[0] * e01) - (x[2] + c[1] * e23)
        assert abs(r @ e01) < 1e-12 and abs(r @ e23) < 1e-12
        assert np.all(np.abs(B.T @ r) < 1e-9 * np.linalg.norm(r) + 1e-15)
    elif kind == "PE":
        e = x[2] - x[1]
        r = x[0] - (x[1] + c[0] * e)
        assert abs(r @ e) < 1e-12
        assert abs(B[:, 1] @ r) < 1e-12 and abs(B[:, 1] @ e) < 1e-12
    else:
        assert np.all(np.abs(B.T @ (x[1] - x[0])) < 1e-12)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("regime", ["slip", "stick"])
def test_friction_gradient_hessian_fd(oracle, kind, regime):
    rng = np.random.default_rng(2 + KINDS.index(kind))
    x0, code = _geom(kind, rng, 3e-4)
    _, lam, c, B = oracle.friction_lag(x0, code, 1e-6, 1e4)
    eps2 = (1e-3) ** 2
    coef = 0.3
    scale = 3e-3 if regime == "slip" else 2e-4
    x = x0 + rng.normal(size=x0.shape) * scale
    ids, g, H = oracle.friction_grad_hess(x, x0, code, c, B, lam, eps2, coef)
    m = len(ids)
    E = lambda y: coef * oracle.friction_energy(y, x0, code, c, B, lam, eps2)
    h = 1e-7 * (1 if regime == "slip" else 0.1)
    gfd = np.zeros(3 * m)
    Hfd = np.zeros((3 * m, 3 * m))
    for a in range(m):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[ids[a], d] += h
            xm[ids[a], d] -= h
            gfd[3 * a + d] = (E(xp) - E(xm)) / (2 * h)
            gp = oracle.friction_grad_hess(xp, x0, code, c, B, lam, eps2, coef)[1]
            gm = oracle.friction_grad_hess(xm, x0, code, c, B, lam, eps2, coef)[1]
            Hfd[:, 3 * a + d] = (gp - gm) / (2 * h)
    assert np.linalg.norm(g - gfd) < 1e-6 * np.linalg.norm(g) + 1e-12
    assert np.linalg.norm(H - Hfd) < 1e-5 * np.linalg.norm(H)
    np.testing.assert_allclose(H, H.T, atol=1e-12 * np.abs(H).max())
    assert np.linalg.eigvalsh(H).min() > -1e-10 * np.abs(H).max()


@pytest.mark.parametrize("regime", ["slip", "stick"])
def test_ground_friction_fd(oracle, regime):
    rng = np.random.default_rng(9)
    n = np.array([0.0, 1.0, 0.0])
    x0 = np.array([0.1, 4e-4, 0.2])
    lam = oracle.friction_lag_ground(x0, n, 0.0, 1e-6, 1e4)
    eps2, coef = (1e-3) ** 2, 0.4
    x = x0 + rng.normal(size=3) * (3e-3 if regime == "slip" else 2e-4)
    g, H = oracle.friction_ground_grad_hess(x, x0, n, lam, eps2, coef)
    E = lambda y: coef * oracle.friction_ground_energy(y, x0, n, lam, eps2)
    h = 1e-8
    gfd = np.array([(E(x + h * e) - E(x - h * e)) / (2 * h) for e in np.eye(3)])
    Hfd = np.stack([(oracle.friction_ground_grad_hess(x + h * e, x0, n, lam, eps2, coef)[0]
                     - oracle.friction_ground_grad_hess(x - h * e, x0, n, lam, eps2, coef)[0])
                    / (2 * h) for e in np.eye(3)], axis=1)
    assert np.linalg.norm(g - gfd) < 1e-6 * np.linalg.norm(g)
    assert np.linalg.norm(H - Hfd) < 1e-5 * np.linalg.norm(H)
    assert np.all(H[1, :] == 0) and np.all(H[:, 1] == 0)


def test_block_sliding_on_ground_decelerates_at_mu_g(oracle):
    """A cube resting on the ground (after settling) given 1 m/s tangential
    velocity decelerates at mu g under ground friction (the lagged normal
    forces sum to M g dt^2) and keeps its speed without it."""
    from sgipc import scene as S
    out = {}
    for mu in (0.0, 0.4):
        verts, tets, part = S.hex_grid((2, 2, 2), 0.05, origin=(0.0, 0.0006, 0.0))
        sc = S.assemble_bodies([(verts, tets, part, 1e6, 0.3, 1000.0, 0)], dt=0.01, dhat=1e-3,
                               ground_friction_mu=mu, friction_mu=0.0)
        sc.abs_xdelta_tol = 1e-7  # converge Newton: the loose default leaves O(1e-2) m/s error
        w = oracle.World(sc)
        for _ in range(15):
            w.time_step()
        x = w.positions()
        v = np.zeros_like(x)
        v[:, 0] = 1.0
        xt = x + v * sc.dt + np.array(sc.gravity) * sc.dt ** 2
        w.set_state(x, v, xt, w.kappa)
        xs = [w.positions()[:, 0].mean()]
        for _ in range(8):
            w.time_step()
            xs.append(w.positions()[:, 0].mean())
        vel = np.diff(xs) / sc.dt
        out[mu] = vel
    decel = (out[0.4][1] - out[0.4][-1]) / ((len(out[0.4]) - 2) * 0.01)
    assert abs(out[0.0][-1] - 1.0) < 0.02
    assert decel == pytest.approx(0.4 * 9.8, rel=0.02)
