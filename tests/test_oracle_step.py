"""Pins of the oracle's time-step control: ComputeTimeStep's bounds in the
cases that reduce to one of them in closed form, and the closed-box ghost
layer."""
import numpy as np

from simcoflow_amd import cases as C


def test_time_step_closed_forms(oracle_lib):
    O = oracle_lib
    g = C.uniform_grid(16, 8, 2.0, 1.0)
    h = g.dx[0]
    z = np.zeros((g.nj + 2, g.ni + 2))
    Fr, cfl = 0.3, 0.25
    # at rest: the gravity-wave bound 2 cfl dy / sqrt(4 dy / Fr) = cfl sqrt(dy Fr)
    dt = O.time_step(g, z, z, cfl, Fr)
    assert abs(dt - cfl * np.sqrt(h * Fr)) <= 1e-15 * dt
    # a fast uniform stream: cfl dx / |u|
    u = z.copy()
    u[1:-1, :] = -40.0
    assert O.time_step(g, u, z, cfl, Fr) == cfl * h / 40.0
    # a vertical stream without gravity: the gravity-wave bound falls to
    # cfl dy / |v| (4 dy / Fr -> 0), equal to the plain vertical bound
    v = z.copy()
    v[1:-1, 1:-1] = 30.0
    assert O.time_step(g, z, v, cfl, 1e30) == cfl * h / 30.0
    # with gravity the gravity-wave bound is the tighter one
    assert O.time_step(g, z, v, cfl, Fr) < cfl * h / 30.0
    # the inflow face u(0,j) alone
    u = z.copy()
    u[3, 0] = 50.0
    assert O.time_step(g, u, z, cfl, Fr) == cfl * h / 50.0
    # the initial vertical velocity scale Vint / Uref
    assert O.time_step(g, z, z, cfl, Fr, vint=20.0, uref=2.0) == cfl * h / 10.0
    # nothing moves and no gravity: the cap of 1
    assert O.time_step(g, z, z, cfl, 1e30) == 1.0


def test_closed_box_ghost_layer(oracle_lib):
    rng = np.random.default_rng(0)
    u, v, p = (rng.standard_normal((7, 9)) for _ in range(3))
    u0, v0, p0 = u.copy(), v.copy(), p.copy()
    oracle_lib.closed_box_bc(u, v, p)
    ni, nj = 7, 5
    assert np.all(u[1:-1, 0] == 0) and np.all(u[1:-1, ni] == 0) and np.all(v[0, 1:-1] == 0)
    assert np.all(p[1:-1, 0] == p0[1:-1, 1]) and np.all(p[1:-1, -1] == p0[1:-1, -2])
    assert np.all(p[0, 1:-1] == p0[1, 1:-1]) and np.all(p[-1, 1:-1] == 0)
    assert np.all(v[1:-1, 0] == v0[1:-1, 1]) and np.all(v[-1, 1:-1] == v0[-2, 1:-1])
    assert np.all(u[-1, 1:-1] == u[-2, 1:-1])
    # interior untouched
    assert np.array_equal(u[1:-1, 1:ni], u0[1:-1, 1:ni]) and np.array_equal(p[1:-1, 1:-1], p0[1:-1, 1:-1])


def test_oracle_steps_keep_volume_and_rest(oracle_lib):
    """A few oracle steps of a bubble: liquid volume kept (the sweeps are
    conservative up to the clipping of the implicit sweep), the correction
    leaves no mass residual, and the bubble starts to rise."""
    from oracle.step import OracleStepper, gas_centroid
    case = C.case_bubble(32, 64)
    st = OracleStepper(oracle_lib, case, tol=1e-12)
    v0 = st.vof.sum()
    c0 = gas_centroid(case.grid, st.vof, case.geo.vofS)
    for _ in range(6):
        rep = st.step()
        assert rep["mres"] < 1e-9
    assert abs(st.vof.sum() - v0) / v0 < 1e-8
    c1 = gas_centroid(case.grid, st.vof, case.geo.vofS)
    assert c1[1] > c0[1]
