"""Pins of the oracle's FEM / ground terms by closed forms and finite
differences (no GPU)."""
import numpy as np
import pytest


def _fd_pk1(O, F, mu, lam, eps=1e-6):
    G = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            Fp, Fm = F.copy(), F.copy()
            Fp[i, j] += eps
            Fm[i, j] -= eps
            G[i, j] = (O.snk1(Fp, mu, lam)[0] - O.snk1(Fm, mu, lam)[0]) / (2 * eps)
    return G


def _vec(M):  # column-major vec (index i + 3 j)
    return M.T.reshape(9)


def test_snk1_rest_state(oracle):
    psi, P = oracle.snk1(np.eye(3), 3.0, 5.0)
    assert psi == 0.0
    assert np.abs(P).max() == 0.0


def test_snk1_uniaxial_closed_form(oracle):
    mu, lam = 2.0, 7.0
    for s in (0.7, 1.0, 1.3, 2.0):
        F = np.diag([s, 1.0, 1.0])
        psi, P = oracle.snk1(F, mu, lam)
        ls = np.log(s)
        assert psi == pytest.approx(mu / 2 * (s * s - 1) - mu * ls + lam / 2 * ls * ls, rel=1e-14, abs=1e-15)
        # dpsi/ds = mu s - mu/s + lam ln(s)/s
        assert P[0, 0] == pytest.approx(mu * s - mu / s + lam * ls / s, rel=1e-13, abs=1e-14)


@pytest.mark.parametrize("seed", range(4))
def test_snk1_derivatives_fd(oracle, seed):
    rng = np.random.default_rng(seed)
    F = np.eye(3) + 0.25 * rng.normal(size=(3, 3))
    if np.linalg.det(F) <= 0.2:
        F[2] *= -1 if np.linalg.det(F) < 0 else 1
    mu, lam = rng.uniform(0.5, 5), rng.uniform(0.5, 20)
    psi, P = oracle.snk1(F, mu, lam)
    assert np.allclose(P, _fd_pk1(oracle, F, mu, lam), rtol=1e-6, atol=1e-7)
    H = oracle.snk1_hess(F, mu, lam, project=False)
    eps = 1e-6
    Hfd = np.zeros((9, 9))
    for b in range(9):
        j, i = divmod(b, 3)
        Fp, Fm = F.copy(), F.copy()
        Fp[i, j] += eps
        Fm[i, j] -= eps
        Hfd[:, b] = (_vec(oracle.snk1(Fp, mu, lam)[1]) - _vec(oracle.snk1(Fm, mu, lam)[1])) / (2 * eps)
    assert np.allclose(H, Hfd, rtol=1e-6, atol=1e-6 * np.abs(H).max())
    # projection = eigen-clamp of the FD Hessian (numpy eigh, independent)
    w, V = np.linalg.eigh(0.5 * (Hfd + Hfd.T))
    Hp = (V * np.maximum(w, 0)) @ V.T
    Ho = oracle.snk1_hess(F, mu, lam, project=True)
    assert np.allclose(Ho, Hp, rtol=1e-5, atol=1e-6 * np.abs(Hp).max())
    assert np.linalg.eigvalsh(Ho).min() > -1e-9 * np.abs(Ho).max()


@pytest.mark.parametrize("seed", range(3))
def test_tet_element_gradient_and_hessian(oracle, seed):
    rng = np.random.default_rng(10 + seed)
    xr = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.1
    Dm = (xr[1:] - xr[0]).T
    Dminv = np.linalg.inv(Dm)
    vol = abs(np.linalg.det(Dm)) / 6
    x = xr + 0.01 * rng.normal(size=xr.shape)
    tet = np.arange(4, dtype=np.uint32)
    mu, lam, dt = 3e3, 1e4, 0.01
    g, H = oracle.tet_grad_hess(x, tet, Dminv, vol, mu, lam, dt)
    # gradient = d/dx of dt^2 vol psi
    eps = 1e-8
    gfd = np.zeros(12)
    for k in range(12):
        xp, xm = x.copy().ravel(), x.copy().ravel()
        xp[k] += eps
        xm[k] -= eps
        gfd[k] = dt * dt * (oracle.tet_energy(xp, tet, Dminv, vol, mu, lam)
                            - oracle.tet_energy(xm, tet, Dminv, vol, mu, lam)) / (2 * eps)
    assert np.allclose(g, gfd, rtol=1e-5, atol=1e-6 * np.abs(g).max())
    # H = dt^2 vol J^T proj(Hq) J with J = dvec(F)/dx from finite differences of F(x)
    def Fof(xx):
        return (xx.reshape(4, 3)[1:] - xx.reshape(4, 3)[0]).T @ Dminv
    J = np.zeros((9, 12))
    for k in range(12):
        xp = x.copy().ravel()
        xp[k] += 1.0
        J[:, k] = _vec(Fof(xp)) - _vec(Fof(x.ravel()))
    Hq = oracle.snk1_hess(Fof(x.ravel()), mu, lam, project=True)
    Href = dt * dt * vol * J.T @ Hq @ J
    assert np.allclose(H, Href, rtol=1e-9, atol=1e-12 * np.abs(Href).max())
    assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()


def test_lumped_mass_and_lame(oracle):
    from sgipc import scene as S
    sc = S.bar_scene()
    w = oracle.World(sc)
    # total kinetic energy of a uniform unit displacement = 0.5 * rho * V
    w.set_xtilde(sc.x + np.array([1.0, 0.0, 0.0]))
    e = w.energy()
    V = 16 * 4 * 4 * 0.05 ** 3
    assert e[1] == pytest.approx(0.5 * 1000.0 * V, rel=1e-12)


def test_ground_barrier(oracle):
    n = np.array([0.0, 1.0, 0.0])
    dhat2 = 1e-6
    kappa = 3.0
    # zero energy and gradient at d^2 = dhat^2
    x = np.array([0.2, 1e-3, -0.1])
    E, g, H = oracle.ground(x, n, 0.0, dhat2, kappa)
    assert abs(E) < 1e-30 and np.abs(g).max() < 1e-20
    for d in (1e-4, 3e-4, 8e-4):
        x = np.array([0.1, d, 0.3])
        E, g, H = oracle.ground(x, n, 0.0, dhat2, kappa)
        t = d * d - dhat2
        assert E == pytest.approx(-kappa * t * t * np.log(d * d / dhat2), rel=1e-13)
        eps = 1e-9
        Ep = oracle.ground(x + eps * n, n, 0.0, dhat2, kappa)[0]
        Em = oracle.ground(x - eps * n, n, 0.0, dhat2, kappa)[0]
        assert g[1] == pytest.approx((Ep - Em) / (2 * eps), rel=1e-5)
        gp = oracle.ground(x + eps * n, n, 0.0, dhat2, kappa)[1]
        gm = oracle.ground(x - eps * n, n, 0.0, dhat2, kappa)[1]
        assert H[1, 1] == pytest.approx((gp[1] - gm[1]) / (2 * eps), rel=1e-5)
        assert H[0, 0] == 0.0 and H[0, 1] == 0.0
