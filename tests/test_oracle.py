"""CPU checks that pin the FP64 oracle against independent answers.

* closed-form eigenpairs of the constant-coefficient operator with the
  reference's boundary mix (Neumann W/E/S, Dirichlet N: ProjectionP_Mod.f90:55);
* hydrostatic known answer for a two-fluid column: the GFM face density of
  ProjectionP_Mod.f90:409-492 reproduces (g/rho_ref) * int rho dy between
  cell centres, and the correction of ComputePUV_Mod.f90:59-152 brings the
  column to rest;
* symmetry and zero row sums (except the Dirichlet row) of the assembled
  cut-cell operator; scipy's direct solve of the oracle's rows;
* PLIC area fractions (Clsvof_Mod.f90:1371-1404) against closed forms and a
  brute-force polygon clip.
"""
import numpy as np
import pytest

from simcoflow_amd import cases as C


def _uniform_case(ni, nj, lx, ly, levelset):
    g = C.uniform_grid(ni, nj, lx, ly)
    geo = C.build_geometry(g, None)
    itf = C.interface_from_levelset(g, geo, levelset)
    return g, geo, itf


def test_eigenpairs_constant_coefficient(oracle_lib):
    O = oracle_lib
    ni, nj = 12, 9
    g, geo, itf = _uniform_case(ni, nj, 1.5, 0.9, lambda X, Y: (-np.ones(X.shape), np.zeros(X.shape), np.ones(X.shape)))
    dt = 0.01
    Dp = np.full((nj + 2, ni + 2), dt)
    fl = C.Fluids()
    M = O.poisson_matrix(g, geo, O.poisson_coef(g, itf, Dp, Dp, fl))
    hx, hy = g.dx[0], g.dy[0]
    cx, cy = dt * hy / hx, dt * hx / hy
    i = np.arange(ni)[None, :]
    j = np.arange(nj)[:, None]
    for k in range(ni):
        for l in range(nj):
            # Neumann-Neumann in x, Neumann-Dirichlet(ghost) in y
            th = (2 * l + 1) * np.pi / (2 * nj + 1)
            v = np.cos(k * np.pi * (i + 0.5) / ni) * np.cos(th * (j + 0.5))
            lam = cx * (2 - 2 * np.cos(k * np.pi / ni)) + cy * (2 - 2 * np.cos(th))
            Av = O.matr_apply(g, M, v)
            assert np.abs(Av - lam * v).max() <= 1e-12 * max(1.0, np.abs(lam * v).max())


def test_hydrostatic_two_fluid_column(oracle_lib):
    O = oracle_lib
    ni, nj = 4, 40
    H = 2.0
    yh = 0.7371  # interface between cell centres
    g, geo, itf = _uniform_case(ni, nj, 0.2, H, lambda X, Y: (Y - yh, np.zeros(X.shape), np.ones(X.shape)))
    fl = C.Fluids(rho_l=1000.0, rho_g=1.0)
    dt = 1e-3
    Dp = np.full((nj + 2, ni + 2), dt)
    # predictor after gravity: v* = -g dt on interior and top faces, walls closed
    u = np.zeros((nj + 2, ni + 2))
    v = np.full((nj + 2, ni + 2), -fl.g * dt)
    v[0, :] = 0.0
    M = O.poisson_matrix(g, geo, O.poisson_coef(g, itf, Dp, Dp, fl))
    b = O.poisson_rhs(g, geo, u, v)
    x, it, rr = O.pcg_mg_solve(g, M, b, tol=1e-13, max_iter=200)
    # exact: p(y) = g/rho_ref * int_y^{H+dy/2} rho dy (p = 0 at the Dirichlet ghost centre)
    hy = g.dy[0]
    top = H + hy / 2

    def integral(y0):
        yl = min(max(yh, y0), top)
        liquid = max(0.0, yh - y0)
        return (fl.rho_l * liquid + fl.rho_g * (top - yl)) / fl.rho_l

    exact = np.array([fl.g * integral(y) for y in g.yc])
    p = x[:, 0]
    assert np.allclose(x, x[:, :1], rtol=0, atol=1e-9 * exact.max())
    assert np.abs(p - exact).max() <= 1e-9 * exact.max()
    # the correction brings the column to rest
    Pp = O.pressure_scatter_bc(g, geo, x)
    uo, vo, po, mres = O.velocity_correct(g, geo, itf, fl, Dp, Dp, u, v, Pp, u, v, np.zeros_like(u))
    assert np.abs(vo[1:nj, 1:-1]).max() <= 1e-9 * fl.g * dt


def test_cut_cell_operator_symmetric_and_conservative(oracle_lib):
    O = oracle_lib
    case = C.case_dambreak(96, 48)
    u, v, DpU, DpV = C.predictor_fields(case)
    M = O.poisson_matrix(case.grid, case.geo, O.poisson_coef(case.grid, case.itf, DpU, DpV, case.fluids))
    A = O.matr_to_scipy(case.grid, case.geo, M).toarray()
    fluid = case.geo.fluid.ravel()
    Af = A[np.ix_(fluid, fluid)]
    assert np.abs(Af - Af.T).max() <= 1e-14 * np.abs(Af).max()
    # rows sum to the Dirichlet (top) term only; elsewhere to the 1e-24 shift
    rs = (M[0] + M[1] + M[2] + M[3] + M[4])
    assert np.abs(rs[:-1][case.geo.fluid[:-1]]).max() <= 1e-12 * np.abs(M[2]).max()
    assert (rs[-1][case.geo.fluid[-1]] > 0).all()
    # positive definite on the fluid rows
    assert np.linalg.eigvalsh(Af).min() > 0
    assert (~case.geo.fluid).sum() > 0 and (case.geo.aE[1:-1, 1:-1] < 1).sum() > 0


@pytest.mark.parametrize("mk", [lambda: C.case_bubble(32, 64), lambda: C.case_swarm(96, 96, seed=5),
                                lambda: C.case_waves(96, 48, seed=1)])
def test_oracle_solver_vs_direct(oracle_lib, mk):
    import scipy.sparse.linalg as spl
    O = oracle_lib
    case = mk()
    u, v, DpU, DpV = C.predictor_fields(case)
    g = case.grid
    M = O.poisson_matrix(g, case.geo, O.poisson_coef(g, case.itf, DpU, DpV, case.fluids))
    b = O.poisson_rhs(g, case.geo, u, v)
    A = O.matr_to_scipy(g, case.geo, M)
    xd = spl.spsolve(A.tocsc(), b.ravel()).reshape(g.shape)
    x, it, rr = O.pcg_mg_solve(g, M, b, tol=1e-11, max_iter=300)
    assert it > 0 and rr <= 1e-11
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) <= 1e-8
    assert O.max_row_residual(g, case.geo, M, x, b) <= 1e-10


def _clip_area(nx, ny, d, hx, hy, n=400):
    xs = (np.arange(n) + 0.5) / n * hx - hx / 2
    ys = (np.arange(n) + 0.5) / n * hy - hy / 2
    X, Y = np.meshgrid(xs, ys)
    return np.mean(nx * X + ny * Y + d < 0)


def test_frac_known_answers():
    hx, hy = 0.3, 0.2
    assert C.frac(0.0, 1.0, -1.0, hx, hy) == 1.0          # all below
    assert C.frac(0.0, 1.0, 1.0, hx, hy) == 0.0           # all above
    assert np.isclose(C.frac(0.0, 1.0, 0.05, hx, hy), 0.25)   # horizontal line y = -0.05
    assert np.isclose(C.frac(1.0, 0.0, -0.075, hx, hy), 0.75)  # vertical line x = 0.075
    s = np.sqrt(0.5)
    assert np.isclose(C.frac(s, s, 0.0, hx, hx), 0.5)          # diagonal of a square
    rng = np.random.default_rng(0)
    for _ in range(40):
        a = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(a), np.sin(a)
        d = rng.uniform(-0.2, 0.2)
        assert abs(C.frac(nx, ny, d, hx, hy) - _clip_area(nx, ny, d, hx, hy)) < 5e-3


def test_box_solid_apertures_known():
    # a solid box covering x in [0.3h, 2.3h]*... : the cut column has
    # aperture 0.7 on its N/S faces and vofS 0.3 (times the covered height)
    n = 12
    g = C.uniform_grid(n, n, 1.0, 1.0)
    h = 1.0 / n
    x0, x1, y0, y1 = 4.3 * h, 7.0 * h, 4.0 * h, 8.0 * h

    def solid(X, Y):
        return C.sdf_box(X, Y, (x0 + x1) / 2, (y0 + y1) / 2, (x1 - x0) / 2, (y1 - y0) / 2)
    geo = C.build_geometry(g, solid)
    # column i=4 (0-based) is 70% solid (x from 4.3h to 5h) in rows 5..6
    assert np.allclose(geo.vofS[5:7, 4], 0.7)
    assert np.allclose(geo.vofS[5:7, 5:7], 1.0)
    # the cut cells' E faces touch solid cells: closed
    assert np.allclose(geo.aE[1 + 5:1 + 7, 1 + 4], 0.0)
    # W faces fully open, N face between the two cut cells 30% open
    assert np.allclose(geo.aW[1 + 5:1 + 7, 1 + 4], 1.0)
    assert np.isclose(geo.aN[1 + 5, 1 + 4], 0.3)
    assert np.isclose(geo.aS[1 + 6, 1 + 4], 0.3)
    # fluid cells away from the box are untouched
    assert geo.vofS[0, 0] == 0.0 and geo.aE[1, 1] == 1.0


def decoupled_case(n=8):
    """A P cell left fluid with all four faces closed: its west and south
    neighbours are solid (closed before it is visited), its east and north
    neighbours are solid (vofS > 1 - epsi) but share an open sliver with it
    when it is visited, and close those faces afterwards (Cell_Geo_Cal,
    Cutcell_Mod.f90:214-243).  The cell itself has a small solid corner, so
    its wall length and normal are non-zero.  Returns (grid, Geometry, the
    cell (j, i))."""
    from simcoflow_amd import cutcell as CC
    g = C.uniform_grid(n, n, 1.0, 1.0)
    h = 1.0 / n
    phi = np.full((n, n), 1.0)
    nx = np.zeros((n, n))
    ny = np.ones((n, n))
    s = 0.7071067811865476
    j0, i0 = 4, 3
    phi[j0, i0 - 1] = -1.0
    phi[j0 - 1, i0] = -1.0
    nx[j0, i0 + 1], ny[j0, i0 + 1], phi[j0, i0 + 1] = -s, s, -(s * 0.5 * h + s * 0.5 * h) + s * 0.03 * h
    nx[j0 + 1, i0], ny[j0 + 1, i0], phi[j0 + 1, i0] = s, -s, -(s * 0.5 * h + s * 0.5 * h) + s * 0.03 * h
    nx[j0, i0], ny[j0, i0], phi[j0, i0] = s, s, -(s * 0.5 * h + s * 0.5 * h) + s * 0.2 * h
    far = (np.full((n, n), 1e3), np.zeros((n, n)), np.ones((n, n)))
    cc = CC.cut_cell_geometry(g, {"P": (phi, nx, ny), "U": far, "V": far})
    return g, C._geometry_from(cc), (j0, i0)


def test_decoupled_fluid_row(oracle_lib):
    """The cell's row is the reference's D = +1e-24 with no coupling
    (ProjectionP_Mod.f90:215-227); with a moving wall its right-hand side
    vb*nyS*WlLh is non-zero, and the solution is b / D there -- what a
    direct solve of the reference's matrix gives -- and the direct solution
    elsewhere."""
    import scipy.sparse.linalg as spl
    O = oracle_lib
    g, geo, (j0, i0) = decoupled_case()
    assert geo.vofS[j0, i0] < 1.0 - 1e-3
    for a in (geo.aE, geo.aW, geo.aN, geo.aS):
        assert a[j0 + 1, i0 + 1] == 0.0
    itf = C.Interface(vof=np.ascontiguousarray(1.0 - geo.vofS), phi=np.full(g.shape, -1.0), nx=np.zeros(g.shape),
                      ny=np.ones(g.shape), vofS=geo.vofS)
    n = g.ni
    rng = np.random.default_rng(4)
    Dp = 0.01 * (1.0 + 0.1 * rng.uniform(size=(n + 2, n + 2)))
    u = 0.1 * rng.standard_normal((n + 2, n + 2))
    v = 0.1 * rng.standard_normal((n + 2, n + 2))
    u[:, 0] = u[:, n] = 0.0
    v[0, :] = 0.0
    M = O.poisson_matrix(g, geo, O.poisson_coef(g, itf, Dp, Dp, C.Fluids()))
    b = O.poisson_rhs(g, geo, u, v, vb=0.2)
    assert M[2, j0, i0] == 1e-24 and not np.any(M[[0, 1, 3, 4], j0, i0])
    assert b[j0, i0] != 0.0
    x, it, rr = O.pcg_mg_solve(g, M, b, tol=1e-12)
    A = O.matr_to_scipy(g, geo, M)
    xd = spl.spsolve(A.tocsc(), b.ravel()).reshape(g.shape)
    assert it > 0 and np.all(np.isfinite(x))
    assert x[j0, i0] == b[j0, i0] / 1e-24
    assert xd[j0, i0] == pytest.approx(x[j0, i0], rel=1e-12)
    mask = np.ones(g.shape, bool)
    mask[j0, i0] = False
    assert np.linalg.norm((x - xd)[mask]) <= 1e-10 * np.linalg.norm(xd[mask])
    assert O.max_row_residual(g, geo, M, x, b) <= 1e-10
