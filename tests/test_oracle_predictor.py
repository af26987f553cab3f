"""Pins of the CPU oracle of the momentum predictor (PredictorUV_mod.f90)
by closed forms that do not restate its formulas:

* a uniform stream in one fluid is a steady state of the convective and
  viscous terms: the predicted U velocity is the stream itself;
* a fluid at rest gains exactly the gravity increment -g dt in V (away from
  the bottom wall and the cut top row), and Dp = dx dy / (aP - sum a_nb)
  is dt wherever there is no wall friction;
* plane Couette shear u = a y is steady in the interior (the viscous fluxes
  of a linear profile cancel);
* the Adams-Bashforth history enters the right-hand side as exactly 0.5
  Flux_n1;
* the two-phase, cut-cell and moving-wall paths are exercised for finite,
  deterministic output.

The MUSCL/kinetic flux details, the mixing weights and the cut-cell
interpolation branches are parity-unpinned: the oracle follows the source
line by line there.
"""
import numpy as np
import pytest

from simcoflow_amd import cases as C

import predictor_cases as P


def _single_phase(ni=32, nj=24, vof=1.0):
    g = C.uniform_grid(ni, nj, 1.0, nj / ni)
    geo = C.build_geometry(g, None)
    U, V = C.staggered_geometry(g, geo, None)
    z = np.zeros(g.shape)
    pcell = dict(vof=np.full(g.shape, vof), phi=np.full(g.shape, -1.0 if vof > 0.5 else 1.0), nx=z, ny=z + 1.0,
                 vofS=geo.vofS, nxS=geo.nxS, nyS=geo.nyS, phiS=geo.phiS)
    uv = np.full(g.shape, vof)
    return g, geo, U, V, pcell, uv


def _run(O, g, U, V, pcell, uv, u, v, phys, dt, first=False, fluxn1=None):
    d = dict(grid=g, U=U, V=V, uvo=uv, uvn=uv, vvo=uv, vvn=uv, pcell=pcell, u=u, v=v, phys=phys, dt=dt,
             fluxn1=np.zeros((2, g.nj + 2, g.ni + 2)) if fluxn1 is None else fluxn1,
             DpU=np.zeros((g.nj + 2, g.ni + 2)), DpV=np.zeros((g.nj + 2, g.ni + 2)),
             upred=np.zeros((g.nj + 2, g.ni + 2)), vpred=np.zeros((g.nj + 2, g.ni + 2)))
    return P.oracle_predictor(O, d, first_step=first)


def test_uniform_stream_is_steady(oracle_lib):
    g, geo, U, V, pcell, uv = _single_phase()
    u = np.full((g.nj + 2, g.ni + 2), 0.37)
    v = np.zeros_like(u)
    phys = P.phys_of(C.Fluids(), Fr=1e30)   # no gravity
    r = _run(oracle_lib, g, U, V, pcell, uv, u, v, phys, dt=0.01)
    up = r["upred"][1:-1, 1:-1][:, :-1]
    assert np.abs(up - 0.37).max() < 1e-13
    assert np.abs(r["vpred"][1:-1, 1:-1][:-1, :]).max() < 1e-13


def test_rest_gains_gravity_and_dp_is_dt(oracle_lib):
    g, geo, U, V, pcell, uv = _single_phase()
    u = np.zeros((g.nj + 2, g.ni + 2))
    v = np.zeros_like(u)
    fl = C.Fluids()
    phys = P.phys_of(fl)
    dt = 0.004
    r = _run(oracle_lib, g, U, V, pcell, uv, u, v, phys, dt)
    vp = r["vpred"][1:-1, 1:-1]
    # 1/Fr^2 = g with Uref = Lref = 1; away from the bottom wall (v = 0 in
    # the ghost row), the side walls and the cut top row the increment is
    # uniform
    assert np.abs(vp[3:-3, 3:-3] / (-fl.g * dt) - 1.0).max() < 1e-12
    assert np.abs(r["upred"][1:-1, 1:-1][:, :-1]).max() < 1e-15
    DU = r["DpU"][1:-1, 1:-1][:, :-1]
    DV = r["DpV"][1:-1, 1:-1][:-1, :]
    assert np.abs(DU - dt).max() < 1e-15 and np.abs(DV - dt).max() < 1e-15


def test_couette_shear_is_steady(oracle_lib):
    g, geo, U, V, pcell, uv = _single_phase(32, 32)
    yc = np.concatenate([[g.yc[0] - g.dy[0]], g.yc, [g.yc[-1] + g.dy[-1]]])
    u = np.ascontiguousarray(np.broadcast_to(0.8 * yc[:, None], (g.nj + 2, g.ni + 2)))
    v = np.zeros_like(u)
    phys = P.phys_of(C.Fluids(), Re=200.0, Fr=1e30)
    r = _run(oracle_lib, g, U, V, pcell, uv, u, v, phys, dt=0.002)
    up = r["upred"][1:-1, 1:-1][:, :-1]
    core = np.s_[6:-6, 6:-6]
    assert np.abs(up[core] - u[1:-1, 1:-1][:, :-1][core]).max() < 1e-12


def test_adams_bashforth_history(oracle_lib):
    """FluxDiv = 1.5 F - 0.5 F_n1 (:211-215): the right-hand side moves by
    exactly 0.5 F_n1 when the history changes (no mixing without solids),
    and the Flux_n1 written back is F, independent of the history read."""
    O = oracle_lib
    case = C.case_bubble(32, 64)
    d0 = P.predictor_inputs(case, seed=1)
    d0["fluxn1"] = np.zeros_like(d0["fluxn1"])
    r0 = P.oracle_predictor(O, d0)
    H = r0["fluxn1"]
    r1 = P.oracle_predictor(O, dict(d0, fluxn1=H))
    assert np.array_equal(r1["fluxn1"], H)
    for comp, key, act in ((0, "rhs_u", d0["U"].active), (1, "rhs_v", d0["V"].active)):
        diff = (r1[key] - r0[key])[act == 1]
        want = 0.5 * H[comp][1:-1, 1:-1][act == 1]
        assert np.abs(diff - want).max() <= 1e-12 * max(np.abs(want).max(), 1e-30)


@pytest.mark.parametrize("name,mk", [("dambreak", lambda: C.case_dambreak(64, 32)),
                                     ("waves", lambda: C.case_waves(48, 24, seed=1))])
def test_cut_cell_paths_finite_and_deterministic(oracle_lib, name, mk):
    O = oracle_lib
    case = mk()
    d = P.predictor_inputs(case, seed=2)
    a = P.oracle_predictor(O, d, vb=0.05)
    b = P.oracle_predictor(O, d, vb=0.05)
    for k in ("matr_u", "rhs_u", "matr_v", "rhs_v", "DpU", "DpV", "fluxn1", "upred", "vpred"):
        assert np.all(np.isfinite(a[k])), k
        assert np.array_equal(a[k], b[k]), k
