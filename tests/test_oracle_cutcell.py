"""The cut-cell geometry oracle (oracle/sg_oracle_cutcell.c: Cell_Geo_Cal,
DefineMomentumExchangeCell, NumberExternalCell, NewCellFace and the two
interpolation routines of Cutcell_Mod.f90) pinned by what the geometry
fixes independently of the formulas -- brute-force edge clipping, walls on
grid lines, planar sloped walls measured in global coordinates, the corner
patch's known values, the order of the closing loop and the stale
north-factor read -- and the host builder (simcoflow_amd/cutcell.py)
against it, bit for bit, on 60 shapes on uniform and stretched grids."""
import numpy as np
import pytest

from simcoflow_amd import cases as C
from simcoflow_amd import cutcell as CC

EPSI = 1.0e-3


def _stretched(ni, nj, seed, lx=2.0, ly=1.5):
    rng = np.random.default_rng(seed)
    dx = rng.uniform(0.6, 1.4, ni)
    dy = rng.uniform(0.6, 1.4, nj)
    dx, dy = lx * dx / dx.sum(), ly * dy / dy.sum()
    xn = np.concatenate([[0.0], np.cumsum(dx)])
    yn = np.concatenate([[0.0], np.cumsum(dy)])
    xc, yc = 0.5 * (xn[1:] + xn[:-1]), 0.5 * (yn[1:] + yn[:-1])
    udx = np.append(xc[1:] - xc[:-1], dx[-1])
    vdy = np.append(yc[1:] - yc[:-1], dy[-1])
    return C.Grid(ni, nj, 0.0, 0.0, dx, dy, udx, vdy, xc, yc)


def _halfplane(nx, ny, c):
    """Solid where nx*x + ny*y + c < 0: signed distance positive in the fluid."""
    n = np.hypot(nx, ny)
    nx, ny, c = nx / n, ny / n, c / n

    def f(X, Y):
        return nx * X + ny * Y + c, np.full(X.shape, nx), np.full(X.shape, ny)
    return f, (nx, ny, c)


# ---------------------------------------------------------------------------
def test_edge_geo_cal_against_brute_force_clipping(oracle_lib):
    """Solid fraction of an edge and the centre of its solid part against
    sampling the edge (Cutcell_Mod.f90:302-337)."""
    O = oracle_lib
    rng = np.random.default_rng(0)
    t = (np.arange(200000) + 0.5) / 200000
    checked = 0
    for _ in range(200):
        h = rng.uniform(0.01, 1.0)
        vertical = rng.uniform() < 0.5
        ang = rng.uniform(0, 2 * np.pi)
        nx, ny = np.cos(ang), np.sin(ang)
        d = rng.uniform(-0.6, 0.6) * h
        if vertical:
            p1, p2 = (0.5 * h, 0.5 * h), (0.5 * h, -0.5 * h)
        else:
            p1, p2 = (-0.5 * h, 0.5 * h), (0.5 * h, 0.5 * h)
        a, fx, fy = O.edge_geo_cal(p1, p2, nx, ny, d)
        px = p1[0] + t * (p2[0] - p1[0])
        py = p1[1] + t * (p2[1] - p1[1])
        solid = nx * px + ny * py + d < 0.0
        assert abs(a - solid.mean()) < 1e-4, (a, solid.mean())
        if 0.0 < solid.mean() < 1.0:
            assert abs(fx - px[solid].mean()) < 1e-5 * h
            assert abs(fy - py[solid].mean()) < 1e-5 * h
            checked += 1
        else:
            assert (fx, fy) == (0.5 * (p1[0] + p2[0]), 0.5 * (p1[1] + p2[1]))
    assert checked > 100


@pytest.mark.parametrize("wall", ["floor", "left", "right"])
def test_walls_on_grid_lines_give_unit_factors(oracle_lib, wall):
    """A straight wall on a P-grid line across the whole domain: P cells all
    or nothing, and every face away from the domain's edge cells has AlE =
    AlN = 1 and DAlE = DAlN = 0 (the centroids of a row, or of a column,
    line up along the wall; in the edge cells the small-cell centroid reads
    the zero aperture ghost)."""
    g = C.uniform_grid(32, 24, 2.0, 1.5)
    h = 2.0 / 32
    f = {"floor": _halfplane(0.0, 1.0, -6 * h), "left": _halfplane(1.0, 0.0, -8 * h),
         "right": _halfplane(-1.0, 0.0, 20 * h)}[wall][0]
    r = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, f))
    P = r["P"]
    assert set(np.unique(P["vofS"])) <= {0.0, 1.0}
    for k in ("aE", "aW", "aN", "aS"):
        a = P[k]
        assert np.all(np.minimum(np.abs(a), np.abs(a - 1.0)) < 1e-12), k
    cut = 0
    for k in "UV":
        T = r[k]
        cut += int(((T["vofS"] > 0) & (T["vofS"] < 1)).sum())
        inner = (slice(2, -2), slice(2, -2))
        assert np.all(T["AlE"][inner] == 1.0) and np.all(T["DAlE"][inner] == 0.0), k
        assert np.all(T["AlN"][inner] == 1.0) and np.all(T["DAlN"][inner] == 0.0), k
    assert cut > 0


def _clip_segment_mid(xa, ya, xb, yb, line):
    """Midpoint of the fluid part (n.x + c >= 0) of the segment a-b, its
    fluid fraction."""
    nx, ny, c = line
    la = nx * xa + ny * ya + c
    lb = nx * xb + ny * yb + c
    if la >= 0 and lb >= 0:
        return 0.5 * (xa + xb), 0.5 * (ya + yb), 1.0
    if la < 0 and lb < 0:
        return None
    t = la / (la - lb)
    xm, ym = xa + t * (xb - xa), ya + t * (yb - ya)
    if la >= 0:
        return 0.5 * (xa + xm), 0.5 * (ya + ym), t
    return 0.5 * (xm + xb), 0.5 * (ym + yb), 1.0 - t


def test_sloped_wall_face_centres_and_u_centroids(oracle_lib):
    """A planar sloped wall on cells larger than 2*epsi: the fluid face
    centre of each cut vertical P face (both sides averaged) and the U-cell
    centroid taken from it (NewCellFace :414-417) are the midpoint of the
    face's fluid segment."""
    g = C.uniform_grid(20, 16, 2.0, 1.6)
    f, line = _halfplane(-0.35, 1.0, -0.55)
    r = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, f))
    P, U = r["P"], r["U"]
    checked = 0
    for j in range(1, g.nj - 1):
        for i in range(1, g.ni - 2):
            xf = g.xc[i] + 0.5 * g.dx[i]
            m = _clip_segment_mid(xf, g.yc[j] - 0.5 * g.dy[j], xf, g.yc[j] + 0.5 * g.dy[j], line)
            if m is None or not (0.02 < m[2] < 0.98):
                continue
            assert abs(g.xc[i] + P["fcex"][j + 1, i + 1] - m[0]) < 1e-12
            assert abs(g.yc[j] + P["fcey"][j + 1, i + 1] - m[1]) < 1e-12
            if U["moex"][j, i] != 1 and U["vofS"][j, i] > EPSI:
                assert abs(g.xc[i] + 0.5 * g.dx[i] + U["ccx"][j, i] - m[0]) < 1e-12
                assert abs(g.yc[j] + U["ccy"][j, i] - m[1]) < 1e-12
            checked += 1
    assert checked >= 10


def test_fine_cells_face_centre_offsets_vanish(oracle_lib):
    """Cutcell_Mod.f90:124 compares the raw face-centre offset, a length,
    with epsi = 1e-3: on cells smaller than 2e-3 every cut vertical face's
    centre sits at the face midpoint."""
    g = C.uniform_grid(64, 64, 0.1, 0.1)
    f, _ = _halfplane(-0.4, 1.0, -0.05)
    r = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, f))
    assert np.all(r["P"]["fcey"][1:-1, 1:-1] == 0.0)
    assert np.all(r["P"]["fcnx"][1:-1, 1:-1] == 0.0)


@pytest.mark.parametrize("slope", [0.35, -0.6, 2.5])
def test_sloped_wall_interpolation_factors(oracle_lib, slope):
    """EastFaceInterpolationInf / NorthFaceInterpolationInf on a planar
    wall: both halves of IntPointDist and FaceCenterDist are the distance
    of one point to the same line, so AlE = |d(face centre)| / |d(point)|
    and AlN = |d(point)| / |d(face centre)| (the reference's north factor is
    the inverse, Cutcell_Mod.f90:577,661), clamped at 2; DAlE = Sy nyS /
    (SxE |d(point)|), DAlN = Sx nxS / (SyN |d(point)|).  The distances are
    measured here from the line's global equation."""
    g = C.uniform_grid(40, 32, 2.0, 1.6)
    f, (nx, ny, c) = _halfplane(-slope, 1.0, -0.55 + 0.3 * slope)
    r = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, f))
    dist = lambda x, y: nx * x + ny * y + c
    axes = CC.grid_axes(g)
    n_e = n_n = 0
    for k in "UV":
        T = r[k]
        x, y, hx, hy = axes[k]
        for j in range(g.nj):
            for i in range(g.ni - 1):
                if not (T["active"][j, i] and T["active"][j, i + 1]):
                    continue
                eta = T["EtaE"][j + 1, i + 1]
                xa, ya = x[i] + T["ccx"][j, i], y[j] + T["ccy"][j, i]
                xb, yb = x[i + 1] + T["ccx"][j, i + 1], y[j] + T["ccy"][j, i + 1]
                Sy = yb - ya
                if not abs(Sy) > 1e-4 * hy[j] * 1.0001:
                    continue
                dp = abs(dist((1 - eta) * xa + eta * xb, (1 - eta) * ya + eta * yb))
                if dp < 1e-5 * 1.0001:
                    continue
                dfc = abs(dist(x[i] + T["fcex"][j + 1, i + 1], y[j] + T["fcey"][j + 1, i + 1]))
                assert T["AlE"][j + 1, i + 1] == pytest.approx(min(dfc / dp, 2.0), rel=1e-9, abs=1e-12)
                assert T["SxE"][j + 1, i + 1] == pytest.approx(xb - xa, rel=1e-12)
                assert T["DAlE"][j + 1, i + 1] == pytest.approx(Sy * ny / ((xb - xa) * dp), rel=1e-9)
                n_e += 1
        for j in range(g.nj - 1):
            for i in range(g.ni):
                if not (T["active"][j, i] and T["active"][j + 1, i]):
                    continue
                eta = T["EtaN"][j + 1, i + 1]
                xa, ya = x[i] + T["ccx"][j, i], y[j] + T["ccy"][j, i]
                xb, yb = x[i] + T["ccx"][j + 1, i], y[j + 1] + T["ccy"][j + 1, i]
                Sx = xb - xa
                if not abs(Sx) >= 1e-4 * hx[i] * 1.0001:
                    continue
                dp = abs(dist((1 - eta) * xa + eta * xb, (1 - eta) * ya + eta * yb))
                if dp < 1e-5 * 1.0001:
                    continue
                dfc = abs(dist(x[i] + T["fcnx"][j + 1, i + 1], y[j] + T["fcny"][j + 1, i + 1]))
                assert T["AlN"][j + 1, i + 1] == pytest.approx(min(dp / dfc, 2.0), rel=1e-9, abs=1e-12)
                assert T["SyN"][j + 1, i + 1] == pytest.approx(yb - ya, rel=1e-12)
                assert T["DAlN"][j + 1, i + 1] == pytest.approx(Sx * nx / ((yb - ya) * dp), rel=1e-9)
                n_n += 1
    assert n_e >= 5 and n_n >= 5, (n_e, n_n)


def test_corner_patch_known_values(oracle_lib):
    """ICorProb (Cutcell_Mod.f90:259-286): a P cell under a full solid cell
    gets WlLh = dx and the normal (0, -1); a P cell under two cells 60 %
    solid gets the bar's values.  The reference writes the north aperture
    from the default-real literal 0.40 and the south one of the cell above
    from 0.4d0; both are 0.4 here (one aperture per face)."""
    g = C.uniform_grid(16, 16, 1.0, 1.0)
    h = 1.0 / 16
    # a block on a grid line (j = 6 below it) and, further right, a bar
    # 0.6 cells wide standing on a grid line (j = 4 below it)
    block = lambda X, Y: C.sdf_box(X, Y, 4 * h, 10 * h, 2 * h, 3 * h)
    bar = lambda X, Y: C.sdf_box(X, Y, 11 * h + 0.3 * h, 8 * h, 0.3 * h, 3 * h)

    def solid(X, Y):
        a, b = block(X, Y), bar(X, Y)
        m = b[0] < a[0]
        return tuple(np.where(m, q, p) for p, q in zip(a, b))
    r = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, solid))
    P = r["P"]
    for i in (2, 3, 4, 5):
        assert P["vofS"][6, i] == 0.0 and P["vofS"][7, i] == 1.0
        assert P["WlLh"][6, i] == h and P["nyS"][6, i] == -1.0 and P["nxS"][6, i] == 0.0
    assert abs(P["vofS"][5, 11] - 0.6) < EPSI and abs(P["vofS"][6, 11] - 0.6) < EPSI
    assert P["vofS"][4, 11] == 0.0
    assert P["aN"][5, 12] == 0.4 and P["aS"][6, 12] == 0.4
    assert P["fcnx"][5, 12] == -0.3 * h and P["fcny"][5, 12] == 0.5 * h
    assert P["WlLh"][4, 11] == 0.6 * h and P["nyS"][4, 11] == -1.0 and P["nxS"][4, 11] == 0.0
    r0 = oracle_lib.cut_cell_geometry(g, C.solid_planes(g, solid), icorprob=False)
    assert r0["P"]["aN"][5, 12] != 0.4 and r0["P"]["WlLh"][4, 11] != 0.6 * h


def test_closing_loop_order(oracle_lib):
    """Cell_Geo_Cal's closing loop (Cutcell_Mod.f90:214-243) visits cells i
    outer, j inner: a fluid cell whose west neighbour is solid and whose east
    neighbour is solid (vofS > 1 - epsi) but still shares an open sliver of
    face is not shut when it is visited -- the east neighbour closes that
    face only afterwards -- so it stays fluid with both x faces closed.
    Closing every solid cell's faces first would shut it."""
    g = C.uniform_grid(8, 8, 1.0, 1.0)
    h = 1.0 / 8
    ni = nj = 8
    phi = np.full((nj, ni), 1.0)
    nx = np.zeros((nj, ni))
    ny = np.ones((nj, ni))
    # west neighbour (i=2, j=4) all solid
    phi[4, 2] = -1.0
    # east neighbour (i=4, j=4): solid except a corner triangle at its NW
    # corner, 0.0015 of the cell wide along its west edge
    s = 0.7071067811865476
    nx[4, 4], ny[4, 4] = -s, s
    phi[4, 4] = -(s * 0.5 * h + s * 0.5 * h) + s * 0.03 * h
    # row j = 2: the same fluid cell between two all-solid cells is shut
    # (W + E < Roundoff) and becomes solid
    phi[2, 2] = -1.0
    phi[2, 4] = -1.0
    planes = {"P": (phi, nx, ny)}
    far = (np.full((nj, ni), 1e3), np.zeros((nj, ni)), np.ones((nj, ni)))
    planes["U"] = planes["V"] = far
    r = oracle_lib.cut_cell_geometry(g, planes)
    P = r["P"]
    assert P["vofS"][4, 4] == 1.0 and P["vofS"][4, 2] == 1.0
    assert P["vofS"][4, 3] == 0.0                       # stays fluid
    assert P["aW"][5, 4] == 0.0 and P["aE"][5, 4] == 0.0  # x faces shut
    assert P["aN"][5, 4] > 0.0 and P["aS"][5, 4] > 0.0
    assert P["vofS"][2, 3] == 1.0 and P["aN"][3, 4] == 0.0 and P["aS"][3, 4] == 0.0
    assert P["aS"][4, 4] == 0.0 and P["aN"][2, 4] == 0.0  # its neighbours' faces
    # the host builder reaches the same state
    cc = CC.cut_cell_geometry(g, planes)
    assert np.array_equal(cc["P"]["vofS"], P["vofS"])


def test_stale_north_factor_read(oracle_lib):
    """AlE(0,j) = AlN(1,j) (Cutcell_Mod.f90:606) reads the north factor the
    previous call left; the rest of AlE does not depend on it."""
    g = C.uniform_grid(24, 20, 1.2, 1.0)
    solid = lambda X, Y: C.sdf_circle(X, Y, 0.6, 0.5, 0.2)
    planes = C.solid_planes(g, solid)
    prev = {k: np.full((g.nj + 2, g.ni + 2), 1.0) for k in "UV"}
    for k in "UV":
        prev[k][1:-1, 1] = 7.0 + np.arange(g.nj)
    r = oracle_lib.cut_cell_geometry(g, planes, prev_AlN=prev)
    r2 = oracle_lib.cut_cell_geometry(g, planes)
    for k in "UV":
        assert np.array_equal(r[k]["AlE"][1:-1, 0], 7.0 + np.arange(g.nj))
        assert np.array_equal(r2[k]["AlE"][1:-1, 0], r2[k]["AlN"][1:-1, 1])
        assert np.array_equal(r[k]["AlE"][1:-1, 1:], r2[k]["AlE"][1:-1, 1:])
        cc = CC.cut_cell_geometry(g, planes, prev_AlN=prev)
        assert np.array_equal(cc[k]["AlE"], r[k]["AlE"])


def _shapes(rng, g, n):
    lx, ly = float(g.dx.sum()), float(g.dy.sum())
    out = []
    for t in range(n):
        kind = ("box", "circle", "ellipse", "boom")[t % 4]
        cx, cy = rng.uniform(0.3, 0.7) * lx, rng.uniform(0.3, 0.7) * ly
        a, b = rng.uniform(0.03, 0.2) * lx, rng.uniform(0.03, 0.2) * ly
        if kind == "box":
            out.append(lambda X, Y, p=(cx, cy, a, b): C.sdf_box(X, Y, *p))
        elif kind == "circle":
            out.append(lambda X, Y, p=(cx, cy, a): C.sdf_circle(X, Y, *p))
        elif kind == "ellipse":
            out.append(lambda X, Y, p=(cx, cy, a, b): C.sdf_ellipse(X, Y, *p))
        else:
            out.append(lambda X, Y, p=(cx, cy, a, rng.uniform(0.1, 0.6) * a, rng.uniform(0.0, 0.2) * ly):
                       C.sdf_boom(X, Y, *p))
    return out


@pytest.mark.parametrize("grid", ["uniform48", "stretched40x56", "odd37x29"])
def test_host_builder_equals_oracle(oracle_lib, grid):
    """cases.build_geometry / staggered_geometry (simcoflow_amd/cutcell.py)
    against the oracle, every P, U and V array bit for bit, over 20 shapes
    in sequence (each call reading the previous one's north factors), then
    no solid."""
    g = {"uniform48": lambda: C.uniform_grid(48, 48, 1.0, 1.0),
         "stretched40x56": lambda: _stretched(40, 56, 3),
         "odd37x29": lambda: C.uniform_grid(37, 29, 2.0, 1.5)}[grid]()
    rng = np.random.default_rng({"uniform48": 1, "stretched40x56": 2, "odd37x29": 3}[grid])
    prev = None
    moex_small = tilted = 0
    for solid in _shapes(rng, g, 20) + [None]:
        planes = C.solid_planes(g, solid)
        ref = oracle_lib.cut_cell_geometry(g, planes, prev_AlN=prev)
        geo, (U, V) = C.cut_cells(g, solid, prev_AlN=prev)
        for name in C.Geometry.__dataclass_fields__:
            key = {"vofSU": ("U", "vofS"), "vofSV": ("V", "vofS"), "moexU": ("U", "moex"),
                   "moexV": ("V", "moex")}.get(name, ("P", name))
            if name in ("uflag", "vflag"):
                T = ref["UV"[name == "vflag"]]
                want = ((T["posnu"] != -1) & (T["moex"] != 1)).astype(np.int32)
            else:
                want = ref[key[0]][key[1]]
            assert np.array_equal(getattr(geo, name), want), name
        for k, st in (("U", U), ("V", V)):
            for name in st.__dataclass_fields__:
                assert np.array_equal(getattr(st, name), ref[k][name]), (k, name)
            moex_small += int(((st.moex == 1) & (st.vofS < 1 - EPSI)).sum())
            tilted += int((st.DAlE != 0).sum() + (st.DAlN != 0).sum())
        prev = {k: ref[k]["AlN"] for k in "UV"}
    assert moex_small > 0 and tilted > 0
