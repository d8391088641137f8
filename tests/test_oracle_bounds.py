"""Oracle pins for the adaptive barrier stiffness (initKappa, suggestKappa,
upperBoundKappa; GIPC.cu:9746-9831) and the Newton step-size bounds (ground
bound GIPC.cu:8857-8915, CFL bound :8917-8968 and their use in solve_subIP
:10786-10822), by closed forms and by what the formulas mean:

* H_b is the second derivative of the IPC barrier b(d) = -(d - dhat)^2 ln(d/dhat)
  in squared distance d (Li et al. 2020), taken here by mpmath at 50 digits;
* suggestKappa makes the barrier's stiffness along the contact normal at
  d = 1e-16 bbox^2, kappa * 4 d * b''(d), equal minKappaCoef * mean mass with
  minKappaCoef = 1e11 (GIPC.cu:8539); mean mass and bbox are closed forms of the
  bar scene;
* initKappa's kappa is the least-squares minimiser of |g_E + kappa g_c| over the
  two gradients the oracle assembles at kappa = 0 and 1, clamped to
  [suggested, 100 x suggested];
* the ground bound is min(1, 0.8 h / (n . dx)) over approaching vertices;
* the CFL bound is 0.5 sqrt(dhat) / max |dx|.
"""
import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")


def _b(d, dhat):
    return -(d - dhat) ** 2 * mpmath.log(d / dhat)


def _b2(d, dhat):
    """b''(d) by mpmath at 50 significant digits"""
    with mpmath.workdps(50):
        return float(mpmath.diff(lambda t: _b(t, mpmath.mpf(dhat)), mpmath.mpf(d), 2))


@pytest.mark.parametrize("ratio", [1e-10, 1e-4, 0.3, 0.9])
def test_H_b_is_the_barrier_second_derivative(oracle, ratio):
    dhat = 1e-6
    d = ratio * dhat
    got = oracle.barrier_H_b(d, dhat)
    want = _b2(d, dhat)
    assert abs(got - want) <= 1e-12 * abs(want)


def test_suggested_kappa_sets_the_normal_stiffness(oracle):
    from sgipc import scene as S
    sc = S.bar_scene()  # 16x4x4 cells of 0.05 m: 0.8 x 0.2 x 0.2 m, rho = 1000
    w = oracle.World(sc)
    k = w.kappa_info()
    # lumped mass closed form: total mass rho * volume spread over the vertices
    mbar = 1000.0 * 0.8 * 0.2 * 0.2 / sc.n_verts
    assert abs(k["mean_mass"] - mbar) <= 1e-12 * mbar
    assert abs(k["bbox_diag2"] - (0.8 ** 2 + 0.2 ** 2 + 0.2 ** 2)) <= 1e-12
    assert k["dhat2"] == sc.dhat ** 2
    d = 1e-16 * k["bbox_diag2"]
    stiffness = k["suggest"] * 4.0 * d * _b2(d, k["dhat2"])
    assert abs(stiffness - 1e11 * mbar) <= 1e-10 * 1e11 * mbar
    # upperBoundKappa caps at 100 times the suggested value
    assert abs(k["upper"] - 100.0 * k["suggest"]) <= 1e-13 * k["upper"]


def _stack_world(oracle, push):
    from sgipc import scene as S
    sc = S.stack_scene(n=2, cells=3)  # contacts with the ground and between the cubes
    w = oracle.World(sc)
    # an inertial target below the cubes presses them into their contacts
    w.set_xtilde(sc.x - np.array([0.0, push, 0.0]))
    return sc, w


@pytest.mark.parametrize("push", [0.0, 5.0, 200.0])
def test_init_kappa_is_the_least_squares_kappa(oracle, push):
    sc, w = _stack_world(oracle, push)
    w.kappa = 0.0
    g0, _ = w.assemble()
    w.kappa = 1.0
    g1, _ = w.assemble()
    gc = g1 - g0
    assert np.linalg.norm(gc) > 0
    k_ls = -float(g0 @ gc) / float(gc @ gc)
    info = w.kappa_info()
    want = min(max(k_ls, info["suggest"]), info["upper"])
    got = w.init_kappa()
    assert abs(got - want) <= 1e-9 * want
    # the three pushes cover the lower clamp, the least-squares branch and the
    # upper clamp
    branch = {0.0: k_ls < info["suggest"], 5.0: info["suggest"] < k_ls < info["upper"],
              200.0: k_ls > info["upper"]}
    assert branch[push]


def test_ground_bound_closed_form(oracle):
    from sgipc import scene as S
    sc = S.bar_scene()  # bottom face 0.1 m above the ground
    w = oracle.World(sc)
    h = sc.x[:, 1].min()
    ones = np.ones((sc.n_verts, 1))
    # 1 / max over surface vertices of (n.dx / (0.8 h) if approaching, else 1):
    # above 1 when every vertex approaches slowly; the step is then capped at 1
    for s, want in [(1.0, 0.8 * h / 1.0), (0.5, 0.8 * h / 0.5), (0.05, 0.8 * h / 0.05),
                    (-1.0, 1.0)]:
        dx = ones * np.array([0.2, s, -0.1])  # rigid translation: no self candidates
        b = w.ccd_bound(dx)
        assert b["n_candidates"] == 0
        assert abs(b["alpha_ground"] - want) <= 1e-14 * want
        assert b["alpha"] == min(1.0, b["alpha_ground"])
    # one vertex approaching faster than the rest sets the bound
    dx = ones * np.array([0.0, 0.1, 0.0])
    v = int(np.argmin(sc.x[:, 1] + 1e-9 * np.arange(sc.n_verts)))
    dx[v, 1] = 0.4
    b = w.ccd_bound(dx)
    assert abs(b["alpha_ground"] - 0.8 * (sc.x[v, 1]) / 0.4) <= 1e-14


def test_cfl_bound_closed_form(oracle):
    sc, w = _stack_world(oracle, 0.0)
    top = sc.x[:, 1] > sc.x[:, 1].mean()
    for u in (0.01, 0.003):
        dx = np.zeros_like(sc.x)
        dx[top, 1] = u  # the top cube moves down towards the lower one
        b = w.ccd_bound(dx)
        assert b["n_candidates"] > 0
        assert abs(b["max_speed"] - u) <= 1e-15
        assert abs(b["alpha_cfl"] - 0.5 * sc.dhat / u) <= 1e-15
        # the final bound follows solve_subIP's rule
        t = min(1.0, b["alpha_ground"], b["alpha_self"])
        want = min(t, b["alpha_cfl"])
        if t > 2 * b["alpha_cfl"]:
            assert b["alpha_cfl"] <= b["alpha"] <= t
        else:
            assert b["alpha"] == want


def test_bank_inverse_is_unpivoted_gauss_jordan(oracle):
    """The MAS bank inverse restates the reference's in-place Gauss-Jordan
    without pivoting: on an SPD matrix it is the inverse, and on a matrix whose
    first pivot is tiny (where partial pivoting would swap rows) it keeps the
    reference's order: its error then exceeds a pivoted solve's."""
    rng = np.random.default_rng(3)
    Q = rng.normal(size=(48, 48))
    A = Q @ Q.T + 48 * np.eye(48)
    Ai = oracle.inverse_gj(A)
    assert np.abs(Ai @ A - np.eye(48)).max() < 1e-12
    # exact on a diagonal and on a unit lower-triangular matrix
    D = np.diag(np.arange(1.0, 49.0))
    assert np.array_equal(oracle.inverse_gj(D), np.diag(1.0 / np.arange(1.0, 49.0)))
    L = np.eye(4)
    L[1, 0] = 2.0
    L[2, 1] = -3.0
    Li = oracle.inverse_gj(L)
    assert np.array_equal(Li @ L, np.eye(4))
    # a tiny leading pivot is used as is
    B = np.array([[1e-12, 1.0], [1.0, 1.0]])
    Bi = oracle.inverse_gj(B)
    want = np.linalg.inv(B)
    assert np.abs(Bi - want).max() > 0  # rounding of the unpivoted order shows
    assert np.abs(Bi - want).max() < 1e-3
