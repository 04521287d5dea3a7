"""CPU tests that pin the oracle's rigid-body coupling (the dynamics block of
immersed_boundary_Wu, boundary_condition.f90:1694-1719, and rk4,
:3168-3202) with closed forms: RK4 integrates a constant acceleration
exactly, RK4 on a linear system multiplies by the degree-4 Taylor polynomial
of its matrix, the symplectic Euler map of the pitch equation preserves area;
and, over whole oracle runs, that the body state follows the force rows the
run reports and moves the lattice when body_rel_pos(1) reaches 1."""
import numpy as np
import pytest

from lbm_mutiwing_amd import SolverConfig, naca0012_markers
from lbm_mutiwing_amd.config import MOTION_STREAM
from lbm_mutiwing_amd.solver import BODY_REL_FIELDS, Dynamics

F = {name: i for i, name in enumerate(BODY_REL_FIELDS)}


def _steps(oracle_mod, dyn, scale, uu, fx, mom, n, st=None):
    st = np.zeros(11) if st is None else st
    hist = []
    for _ in range(n):
        oracle_mod.body_step(dyn, scale, uu, fx, mom, st)
        hist.append(st.copy())
    return np.array(hist)


@pytest.mark.parametrize("closed,mass", [(True, 4.0), (False, 8.0)])
def test_rk4_constant_force_is_exact(oracle_mod, closed, mass):
    # solid_den 2, area 0.5, scale 0.5: m = 2*0.5/0.25 or 2/0.25, all binary fractions
    dyn = Dynamics(closed_boundary=closed, rk4=True, solid_den=2.0, area=0.5, par1=0.0, par2=0.0)
    fx = 0.5
    a = fx / mass
    h = _steps(oracle_mod, dyn, 0.5, 0.1, fx, 0.0, 12)
    n = np.arange(1, 13, dtype=float)
    assert np.array_equal(h[:, F["a"]], np.full(12, a))
    assert np.array_equal(h[:, F["v"]], a * n)
    assert np.array_equal(h[:, F["s"]], a * n * n / 2)
    assert np.array_equal(h[:, F["pos"]], a * n * n / 2)
    assert np.array_equal(h[1:, F["a_prev"]], np.full(11, a)) and h[0, F["a_prev"]] == 0.0
    assert np.array_equal(h[:, F["v_prev"]], a * (n - 1))
    assert np.array_equal(h[:, F["s_prev"]], a * (n - 1) ** 2 / 2)


def test_euler_constant_force_closed_form(oracle_mod):
    dyn = Dynamics(closed_boundary=True, rk4=False, solid_den=2.0, area=0.5, par1=0.0, par2=0.0)
    a = 0.5 / 4.0
    h = _steps(oracle_mod, dyn, 0.5, 0.1, 0.5, 0.0, 10)
    n = np.arange(1, 11, dtype=float)
    assert np.array_equal(h[:, F["v"]], a * n)
    assert np.array_equal(h[:, F["s"]], a * n * (n - 1) / 2)      # s advances by the previous velocity
    assert np.array_equal(h[:, F["pos"]], a * n * (n - 1) / 2)


def test_rk4_pitch_is_taylor_polynomial_of_linear_system(oracle_mod):
    # theta'' = p1 (theta + alphad) with p1 = par1*UU^2*scale^2 = -0.0025:
    # y = (theta + alphad, omega), y' = A y; one RK4 step is y <- T4(A) y
    uu, scale, alphad = 0.1, 0.5, 0.3
    dyn = Dynamics(rk4=True, solid_den=1.0, area=1.0, par1=-1.0, par2=0.0, alphad=alphad)
    p1 = -1.0 * uu * uu * scale * scale
    A = np.array([[0.0, 1.0], [p1, 0.0]])
    T4 = np.eye(2) + A + A @ A / 2 + A @ A @ A / 6 + A @ A @ A @ A / 24
    st = np.zeros(11)
    st[F["theta"]], st[F["temp_theta"]] = 0.2, -0.01
    y = np.array([0.2 + alphad, -0.01])
    h = _steps(oracle_mod, dyn, scale, uu, 0.0, 123.0, 40, st)
    for k in range(40):
        y = T4 @ y
        assert abs(h[k, F["theta"]] + alphad - y[0]) < 1e-13 and abs(h[k, F["temp_theta"]] - y[1]) < 1e-15
        assert h[k, F["theta_prev"]] == (h[k - 1, F["theta"]] if k else 0.2)
    # and the exact harmonic solution to RK4's global error
    w = np.sqrt(-p1)
    t = 40.0
    exact = (0.2 + alphad) * np.cos(w * t) + (-0.01 / w) * np.sin(w * t)
    assert abs(h[-1, F["theta"]] + alphad - exact) < 2e-7   # 40 steps of local error ~ w^5/120


def test_rk4_pitch_constant_moment_is_exact(oracle_mod):
    # p1 = 0: theta'' = par2*scale^4*mom, a quadratic that RK4 integrates exactly
    dyn = Dynamics(rk4=True, solid_den=1.0, area=1.0, par1=0.0, par2=1.0, alphad=0.7)
    c = 1.0 * 0.5 ** 4 * 0.25
    st = np.zeros(11)
    st[F["theta"]], st[F["temp_theta"]] = 0.5, 0.125
    h = _steps(oracle_mod, dyn, 0.5, 0.1, 0.0, 0.25, 8, st)
    n = np.arange(1, 9, dtype=float)
    assert np.array_equal(h[:, F["temp_theta"]], 0.125 + c * n)
    assert np.array_equal(h[:, F["theta"]], 0.5 + 0.125 * n + c * n * n / 2)


def test_euler_pitch_map_preserves_area(oracle_mod):
    # theta <- theta + omega; omega <- omega + p1 (theta_new + alphad): det 1.
    # Updating omega with the old theta instead would scale areas by 1 - p1.
    dyn = Dynamics(rk4=False, solid_den=1.0, area=1.0, par1=-4.0, par2=0.0, alphad=0.0)
    ends = []
    for y0 in ((1.0, 0.0), (0.0, 1.0)):
        st = np.zeros(11)
        st[F["theta"]], st[F["temp_theta"]] = y0
        h = _steps(oracle_mod, dyn, 0.5, 0.1, 0.0, 0.0, 200, st)
        ends.append((h[-1, F["theta"]], h[-1, F["temp_theta"]]))
    det = ends[0][0] * ends[1][1] - ends[0][1] * ends[1][0]
    assert abs(det - 1.0) < 1e-12


def _tiny_case(wings=1):
    s = (0.0, 1.4, 1.4, 0.0, 1.2, 0.0, 1.2, 0.0, 0.0)
    cfg = SolverConfig(nx=192, ny=96, scale=1 / 24, uu=0.05, re=100.0, s=s, wing_count=wings,
                       markers=naca0012_markers(24), ox=-1.5, oy=0.5 - 0.5 * 95 / 24,
                       motion_source=MOTION_STREAM, name="tiny")
    cfg.wing_dx = 2.0
    return cfg


@pytest.mark.parametrize("wings,rk4,solid_den", [(1, True, 40.0), (2, True, 100.0), (2, False, 100.0)])
def test_oracle_run_body_follows_its_force_rows(oracle_mod, wings, rk4, solid_den):
    """Over a run with the coupling on, the body state equals the update
    replayed from the run's own force rows (total Fx and wing 1's moment,
    un-normalised), the rows' x columns carry the updated displacement, and
    reaching body_rel_pos(1) >= 1 moves the lattice once."""
    cfg = _tiny_case(wings)
    uu, scale, W = cfg.uu, cfg.scale, wings
    # wing 1 starts drifting downstream at 0.02 lattice units per step, half
    # a node from a move; the densities keep the explicit coupling stable
    # (a body lighter than the fluid in its IBM support oscillates and grows)
    dyn = Dynamics(rk4=rk4, solid_den=solid_den, par1=-2.0, par2=50.0)
    o = oracle_mod.OracleSolver(cfg)
    o.reset()
    o.set_dynamics(dyn)
    rel0 = np.zeros((11, W))
    rel0[F["v"], 0], rel0[F["pos"], 0] = 0.02, 0.5
    o.set_body_rel(rel0)
    K = 60
    rows = o.advance(K, want_rows=True).force_rows
    st = rel0[:, 0].copy()
    moved = 0
    for k in range(K):
        fx = rows[k, 1 + 3 * W] * (uu * uu * (1.0 / scale)) / 2.0     # total Fx, raw
        mom = -rows[k, 3] * (uu * uu * (1.0 / (scale * scale))) / 2.0  # wing 1 moment, raw
        oracle_mod.body_step(dyn, scale, uu, fx, mom, st)
        # x column of wing w: wing_x0 + (w-1) dx + sa + body_extra_s * scale (sa = 0 here)
        for w in range(W):
            xcol = rows[k, 1 + 3 * W + 3 + 9 * w]
            assert abs(xcol - (cfg.wing_x0 + w * cfg.wing_dx + st[F["s"]] * scale)) < 1e-12 * max(1.0, abs(xcol))
        if st[F["pos"]] >= 1.0:
            st[F["pos"]] -= 1.0
            moved += 1
    rel = o.body_rel()
    assert np.allclose(rel[:, 0], st, rtol=1e-9, atol=1e-13), (rel[:, 0], st)
    assert np.all(rel[:, 1:] == 0.0)
    assert moved == 1 and o.state()["movenum"] == 1
    assert rel[F["s"], 0] > 1.0


def test_body_rel_round_trip_and_reset(oracle_mod):
    cfg = _tiny_case(2)
    o = oracle_mod.OracleSolver(cfg)
    o.reset()
    rel = np.arange(22, dtype=float).reshape(11, 2) / 8.0
    o.set_body_rel(rel)
    assert np.array_equal(o.body_rel(), rel)
    o.reset()
    assert np.array_equal(o.body_rel(), np.zeros((11, 2)))
    with pytest.raises(Exception):
        o.set_dynamics(Dynamics(solid_den=0.0))
