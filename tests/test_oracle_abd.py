"""Pins of the oracle's affine-body terms (abd_system/details/abd_energy.inl,
abd_jacobi_matrix.h, setup_abd_system_gradient_and_hessian.cu): the shape
energy's closed form, finite differences of its gradient and Hessian, the
Jacobian, the node-space gradient of a world against finite differences of
its incremental potential, the body blocks against their definition, and
rigid motion of affine boxes (free fall, spinning, resting on a soft box)."""
import numpy as np
import pytest


def _rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def _q(p, A):
    return np.concatenate([p, A[0], A[1], A[2]])


def test_shape_energy_closed_form(oracle):
    rng = np.random.default_rng(0)
    R = _rot([1, 2, 3], 0.7)
    e, g, _ = oracle.abd_shape(_q(rng.normal(size=3), R))
    assert e == pytest.approx(0.0, abs=1e-28)
    assert np.abs(g).max() < 1e-14
    for _ in range(3):
        A = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
        e, _, _ = oracle.abd_shape(_q(np.zeros(3), A))
        assert e == pytest.approx(np.linalg.norm(A @ A.T - np.eye(3)) ** 2, rel=1e-13)


@pytest.mark.parametrize("seed", range(4))
def test_shape_derivatives_fd(oracle, seed):
    rng = np.random.default_rng(10 + seed)
    q = _q(rng.normal(size=3), np.eye(3) + 0.3 * rng.normal(size=(3, 3)))
    e, g, H = oracle.abd_shape(q)
    h = 1e-6
    gfd = np.zeros(9)
    Hfd = np.zeros((9, 9))
    for k in range(9):
        d = np.zeros(12)
        d[3 + k] = h
        gfd[k] = (oracle.abd_shape(q + d)[0] - oracle.abd_shape(q - d)[0]) / (2 * h)
        Hfd[:, k] = (oracle.abd_shape(q + d)[1] - oracle.abd_shape(q - d)[1]) / (2 * h)
    np.testing.assert_allclose(g, gfd, rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(H, Hfd, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(H, H.T, atol=1e-13)


def test_jacobian(oracle):
    rng = np.random.default_rng(3)
    xb = rng.normal(size=3)
    J = oracle.abd_jmat(xb)
    p, A = rng.normal(size=3), rng.normal(size=(3, 3))
    np.testing.assert_allclose(J @ _q(p, A), p + A @ xb, rtol=1e-14)


def _falling_box(rot=0.0, omega=None, **kw):
    from sgipc import scene as S
    x0, t0, p0 = S.hex_grid((2, 2, 2), 0.05)
    x0 = x0 - x0.mean(0)
    x0 = x0 @ _rot([0, 1, 0], rot).T + np.array([0.0, 1.0, 0.0])
    sc = S.assemble_bodies([(x0, t0, p0, 1e9, 0.3, 1000.0, 0, True)], dt=0.01, dhat=1e-3,
                           use_ground=False, **kw)
    if omega is not None:
        c = x0.mean(0)
        sc.v = np.cross(omega, x0 - c)
    sc.abs_xdelta_tol = 1e-7
    return sc


def test_free_fall_is_rigid(oracle):
    sc = _falling_box(rot=0.3)
    w = oracle.World(sc)
    assert w.nodes == 4
    x0 = w.positions().copy()
    for n in range(1, 5):
        w.time_step()
        drop = -9.8 * 0.01 ** 2 * n * (n + 1) / 2
        np.testing.assert_allclose(w.positions(), x0 + np.array([0.0, drop, 0.0]), atol=1e-10)
    assert w.energy()[7] == pytest.approx(0.0, abs=1e-20)


def test_spinning_box_stays_rigid(oracle):
    """an affine box spinning at 5 rad/s: the shape energy (kappa = 1e8) keeps
    A a rotation, so every edge length is preserved to 1e-5"""
    sc = _falling_box(omega=np.array([0.0, 5.0, 2.0]), gravity=(0.0, 0.0, 0.0))
    w = oracle.World(sc)
    x0 = w.positions().copy()
    d0 = np.linalg.norm(x0[:, None] - x0[None], axis=2)
    for _ in range(10):
        w.time_step()
    x = w.positions()
    d = np.linalg.norm(x[:, None] - x[None], axis=2)
    assert np.abs(d - d0).max() / d0.max() < 1e-5
    assert np.abs(x - x0).max() > 0.01          # it did turn


def test_node_gradient_is_energy_derivative(oracle):
    """the node gradient is d(incremental potential)/dq for a body in contact
    (kinetic M (q - q~), shape, barrier through J^T) and d/dx for FEM nodes"""
    from sgipc import scene as S
    sc = S.abd_scene(n_abd=2, gap=4e-4)
    sc.kappa = 1e4
    w = oracle.World(sc)
    x = w.positions().copy()
    # move inside the bodies' affine space: perturb q of body 0 (rotation + stretch)
    abd = sc.abd_body
    xb = x[abd == 0] - x[abd == 0].mean(0)
    A = np.eye(3) + np.array([[0.01, 0.003, 0], [0, -0.004, 0.002], [0.001, 0, 0.006]])
    x[abd == 0] = x[abd == 0].mean(0) + 1e-4 + xb @ A.T
    w.set_positions(x)
    w.kappa = 1e4
    g, _ = w.assemble()
    nfem = int((abd < 0).sum())
    gb = g[3 * nfem:3 * nfem + 12]
    xbar0 = sc.x[abd == 0] - sc.x[abd == 0].mean(0)   # body frame fixed at creation
    J = [oracle.abd_jmat(v) for v in xbar0]
    h = 1e-7
    for k in (0, 1, 4, 8, 11):
        def E(s):
            y = x.copy()
            dq = np.zeros(12)
            dq[k] = s
            y[abd == 0] = x[abd == 0] + np.array([Jm @ dq for Jm in J])
            w.set_positions(y)
            return w.energy()[0]
        fd = (E(h) - E(-h)) / (2 * h)
        assert gb[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(gb).max())
    w.set_positions(x)


def test_body_block_is_mass_plus_projected_shape(oracle):
    sc = _falling_box(rot=0.2)
    w = oracle.World(sc)
    x = w.positions()
    c = x.mean(0)   # uniform masses of a regular grid: the mass centre is the mean
    A = np.eye(3) + np.array([[0.02, 0.01, 0], [0, -0.03, 0.01], [0.0, 0, 0.04]])
    w.set_positions(c + (x - c) @ A.T)
    _, H = w.assemble()
    Hd = H.dense()
    # M = sum m J^T J with lumped masses (a regular 2x2x2 grid of equal tets)
    from sgipc import scene as S
    g, _ = w.assemble()
    xbar = x - c
    vol = 0.05 ** 3 / 6
    m = np.zeros(len(x))
    for t in sc.tets:
        m[t] += vol * 1000.0 / 4
    M = sum(mi * oracle.abd_jmat(xb).T @ oracle.abd_jmat(xb) for mi, xb in zip(m, xbar))
    q = np.concatenate([c, A[0], A[1], A[2]])
    _, _, Hs = oracle.abd_shape(q)
    Hs = Hs * 1e8 * 0.1 ** 3 * 0.01 ** 2
    wv, V = np.linalg.eigh(Hs)
    Hp = (V * np.maximum(wv, 0)) @ V.T
    want = M.copy()
    want[3:, 3:] += Hp
    np.testing.assert_allclose(Hd, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_abd_stack_on_soft_box_rests(oracle):
    """two affine boxes stacked on a soft FEM box settle: no interpenetration
    (the bodies keep their gaps positive) and the affine boxes stay rigid"""
    from sgipc import scene as S
    sc = S.abd_scene(n_abd=2)
    w = oracle.World(sc)
    x0 = w.positions().copy()
    for _ in range(12):
        w.time_step()
    x = w.positions()
    abd = sc.abd_body
    for b in (0, 1):
        d0 = np.linalg.norm(x0[abd == b][:, None] - x0[abd == b][None], axis=2)
        d = np.linalg.norm(x[abd == b][:, None] - x[abd == b][None], axis=2)
        assert np.abs(d - d0).max() < 1e-6
    fem = abd < 0
    assert x[abd == 0, 1].min() > x[fem, 1].max() - 1e-3   # resting on the soft box
    assert x[abd == 1, 1].min() > x[abd == 0, 1].max()
    assert x[:, 1].min() > 0.0


def test_block_jacobi_inverts_body_blocks_whole(oracle):
    """Block-Jacobi on an affine-body system (DiagPreconditioner on the FEM
    nodes, ABDPreconditioner on each body, abd_preconditioner.cu:15-38): z on a
    body's 4 nodes = inverse of its 12x12 diagonal block times r (numpy), on
    FEM nodes the 3x3 inverse; and it is a better preconditioner for the body
    than its 3x3 node blocks would be"""
    from sgipc import scene as S
    sc = S.abd_scene(n_abd=2, gap=4e-4)
    sc.kappa = 1e4
    w = oracle.World(sc)
    g, A = w.assemble()
    assert A.nbody == 2
    D = A.dense()
    rng = np.random.default_rng(0)
    r = rng.normal(size=3 * A.nb)
    z = oracle.block_jacobi(A, r)
    nb3 = A.nb - 8
    for i in range(nb3):
        blk = D[3 * i:3 * i + 3, 3 * i:3 * i + 3]
        np.testing.assert_allclose(z[3 * i:3 * i + 3], np.linalg.solve(blk, r[3 * i:3 * i + 3]),
                                   rtol=1e-10, atol=1e-14 * np.abs(z).max())
    for b in range(2):
        s = slice(3 * (nb3 + 4 * b), 3 * (nb3 + 4 * b + 4))
        blk = D[s, s]
        np.testing.assert_allclose(z[s], np.linalg.solve(blk, r[s]), rtol=1e-8,
                                   atol=1e-12 * np.abs(z[s]).max())
    x12, it12, _, _ = oracle.pcg(A, g, None, 0, rel_tol=1e-6)
    A.nbody = 0  # the same system preconditioned by 3x3 node blocks only
    x3, it3, _, _ = oracle.pcg(A, g, None, 0, rel_tol=1e-6)
    assert it12 < it3


def _jac(xbar):
    """ABDJacobi::to_mat (abd_jacobi_matrix.inl:32-43): x = t + (a1.xbar, a2.xbar, a3.xbar)"""
    J = np.zeros((3, 12))
    J[:, :3] = np.eye(3)
    for i in range(3):
        J[i, 3 + 3 * i:6 + 3 * i] = xbar
    return J


def test_motor_target_is_the_reference_matrix_formula(oracle):
    """q_p = J^-1 J_delta q~ + q~ with J, J_delta stacked from the points 0,
    e_x, e_y, e_z and their displacements under R_x(theta)
    (setup_abd_system_gradient_and_hessian.cu:264-307), rows a2 and a3
    normalised, restated with numpy matrices"""
    rng = np.random.default_rng(4)
    for theta in (-0.02 * 0.01, 0.3, -1.1):
        qt = _q(rng.normal(size=3), _rot(rng.normal(size=3), 0.8) @ (np.eye(3) + 0.05 *
                                                                   rng.normal(size=(3, 3))))
        e = np.eye(3)
        R = _rot([1, 0, 0], theta)
        J = np.vstack([_jac(np.zeros(3)), _jac(e[0]), _jac(e[1]), _jac(e[2])])
        Jd = np.vstack([_jac(np.zeros(3)), _jac(np.zeros(3)), _jac(R @ e[1] - e[1]),
                        _jac(R @ e[2] - e[2])])
        ref = np.linalg.inv(J) @ Jd @ qt + qt
        for s in (slice(6, 9), slice(9, 12)):
            ref[s] /= np.linalg.norm(ref[s])
        got = oracle.abd_motor_target(qt, theta)
        np.testing.assert_allclose(got[3:], ref[3:], rtol=1e-12, atol=1e-14)


def _motor_scene():
    from sgipc import scene as S
    sc = S.abd_scene(n_abd=1, fem_base=False, gravity=(0.0, 0.0, 0.0), use_ground=False,
                     motor_speed=-2.0, motor_strength=200.0)
    sc.btype = sc.btype.copy()
    sc.btype[sc.abd_body == 0] = 2
    return sc


def test_motor_gradient_is_energy_derivative(oracle):
    sc = _motor_scene()
    w = oracle.World(sc)
    w.time_step()  # sets the motor target from the step prediction
    x = w.positions().copy()
    g, _ = w.assemble()
    abd = sc.abd_body
    xbar0 = sc.x[abd == 0] - sc.x[abd == 0].mean(0)
    J = [oracle.abd_jmat(v) for v in xbar0]
    h = 1e-7
    for k in (0, 4, 6, 8, 10, 11):
        def E(s):
            y = x.copy()
            dq = np.zeros(12)
            dq[k] = s
            y[abd == 0] = x[abd == 0] + np.array([Jm @ dq for Jm in J])
            w.set_positions(y)
            return w.energy()[0]
        fd = (E(h) - E(-h)) / (2 * h)
        assert g[k] == pytest.approx(fd, rel=1e-5, abs=1e-9 * np.abs(g).max())
    w.set_positions(x)


def test_motor_turns_the_body_about_its_x_axis(oracle):
    """a free motor box (no gravity, no contact) turns about its body x axis
    and stays rigid.  The target rotates the step prediction q~ (which carries
    the current angular velocity) by theta = speed * dt, so each step adds
    theta to the angular step: after n steps the angle is theta n (n + 1) / 2"""
    sc = _motor_scene()
    w = oracle.World(sc)
    abd = sc.abd_body
    x0 = w.positions()[abd == 0]
    th = sc.motor_speed * sc.dt
    for n in range(1, 7):
        w.time_step()
        x = w.positions()[abd == 0]
        A, *_ = np.linalg.lstsq(x0 - x0.mean(0), x - x.mean(0), rcond=None)
        A = A.T
        assert np.linalg.norm(A @ A.T - np.eye(3)) < 1e-3 and abs(A[0, 0] - 1) < 1e-3
        ang = np.arctan2(A[2, 1], A[1, 1])
        assert ang == pytest.approx(th * n * (n + 1) / 2, rel=0.05)
