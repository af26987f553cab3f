"""Case set-up for the projection / interface-transport path.

The reference hard-codes one oil-boom case in ``Main.f90:146-206``; this module
builds the inputs that path consumes for the benchmark workloads instead:

* the tensor-product grid of ``Mesh_Mod.f90:54-175`` (P, U and V spacings);
* the cut-cell solid geometry from a solid signed-distance field: solid volume
  fraction through the PLIC ``frac`` of ``Clsvof_Mod.f90:1371-1404`` (as
  ``Initial_Clsvof`` does at ``Clsvof_Mod.f90:273-277``), edge apertures,
  wall length, harmonic face averaging and solid closure as in
  ``Cell_Geo_Cal`` (``Cutcell_Mod.f90:34-300``), momentum-exchange cells
  (``Cutcell_Mod.f90:360-397``) and the fluid numbering test
  (``Cutcell_Mod.f90:340-358``);
* liquid volume fraction / level set of the interface (gas where phi > 0).

Arrays use the reference's Fortran layout: a cell array has numpy shape
(nj, ni) in C order (x fastest); a ghosted array has shape (nj+2, ni+2) with
the (0,0) ghost at [0, 0].
"""
from dataclasses import dataclass, field
import numpy as np

EPSI = 1.0e-3      # StateVariables.f90:28
EPSIF = 1.0e-2     # StateVariables.f90:28
VOFEPS = 1.0e-14   # Clsvof_Mod.f90:10
ROUNDOFF = 1.0e-24  # Cutcell_Mod.f90:45


@dataclass
class Grid:
    ni: int
    nj: int
    x0: float
    y0: float
    dx: np.ndarray   # (ni,) P-cell widths
    dy: np.ndarray   # (nj,) P-cell heights
    udx: np.ndarray  # (ni,) U-cell widths  (Mesh_Mod.f90:154-158)
    vdy: np.ndarray  # (nj,) V-cell heights (Mesh_Mod.f90:167-171)
    xc: np.ndarray   # (ni,) P-cell centres
    yc: np.ndarray   # (nj,) P-cell centres

    @property
    def shape(self):
        return (self.nj, self.ni)

    @property
    def ncells(self):
        return self.ni * self.nj


def uniform_grid(ni, nj, lx, ly, x0=0.0, y0=0.0):
    """Uniform grid as Initial_Grid builds it with NonUniformMesh=0
    (Mesh_Mod.f90:93-96,123-135) and the staggered spacings of
    InitialUVGrid (Mesh_Mod.f90:139-175)."""
    hx = lx / ni
    hy = ly / nj
    xn = x0 + hx * np.arange(ni + 1)
    yn = y0 + hy * np.arange(nj + 1)
    xc = 0.5 * (xn[1:] + xn[:-1])
    yc = 0.5 * (yn[1:] + yn[:-1])
    dx = xn[1:] - xn[:-1]
    dy = yn[1:] - yn[:-1]
    udx = np.empty(ni)
    udx[:-1] = xc[1:] - xc[:-1]
    udx[-1] = dx[-1]
    vdy = np.empty(nj)
    vdy[:-1] = yc[1:] - yc[:-1]
    vdy[-1] = dy[-1]
    return Grid(ni, nj, x0, y0, dx, dy, udx, vdy, xc, yc)


@dataclass
class Fluids:
    """Fluid constants; the reference hard-codes water/air at 20 C
    (StateVariables.f90:26-27) with Roref = row (Main.f90:148)."""
    rho_l: float = 1000.0
    rho_g: float = 1.0
    mu_l: float = 1.0e-3
    mu_g: float = 1.0e-5
    g: float = 9.81

    @property
    def rho_ref(self):
        return self.rho_l


# ---------------------------------------------------------------------------
# PLIC area fraction (Clsvof_Mod.f90:1371-1404), vectorised.
def frac(nx, ny, d, dx, dy):
    """Area fraction of a dx*dy cell (centred on 0) where nx*x+ny*y+d < 0."""
    nx, ny, d, dx, dy = np.broadcast_arrays(*(np.asarray(a, dtype=np.float64) for a in (nx, ny, d, dx, dy)))
    tnx = np.abs(nx)
    tny = np.abs(ny)
    out = np.empty(nx.shape)
    half = 0.5 * (tnx * dx + tny * dy)
    full = d + half <= 0.0
    empty = (~full) & (d - half >= 0.0)
    rest = ~(full | empty)
    out[full] = 1.0
    out[empty] = 0.0
    vx = rest & (tnx <= 1e-20)
    out[vx] = 0.5 - d[vx] / dy[vx]
    vy = rest & ~vx & (tny <= 1e-20)
    out[vy] = 0.5 - d[vy] / dx[vy]
    gen = rest & ~vx & ~vy
    if np.any(gen):
        a, b, dd, hx, hy = tnx[gen], tny[gen], d[gen], dx[gen], dy[gen]
        xx = (0.5 * b * hy - dd) / hx / a + 0.5
        yy = (0.5 * a * hx - dd) / hy / b + 0.5
        total = 0.5 * xx * yy
        top = np.where(yy > 1.0, ((yy - 1.0) / np.where(yy > 0, yy, 1.0)) ** 2, 0.0)
        right = np.where(xx > 1.0, ((xx - 1.0) / np.where(xx > 0, xx, 1.0)) ** 2, 0.0)
        out[gen] = total * (1.0 - top - right)
    return out


# ---------------------------------------------------------------------------
# Signed-distance shapes: value > 0 outside the shape, with unit gradient.
def sdf_box(X, Y, cx, cy, hw, hh):
    qx = np.abs(X - cx) - hw
    qy = np.abs(Y - cy) - hh
    ox = np.maximum(qx, 0.0)
    oy = np.maximum(qy, 0.0)
    outside = np.sqrt(ox * ox + oy * oy)
    inside = np.minimum(np.maximum(qx, qy), 0.0)
    phi = outside + inside
    sx = np.sign(X - cx)
    sy = np.sign(Y - cy)
    sx = np.where(sx == 0, 1.0, sx)
    sy = np.where(sy == 0, 1.0, sy)
    out = outside > 0
    nrm = np.where(out, outside, 1.0)
    gx = np.where(out, sx * ox / nrm, np.where(qx >= qy, sx, 0.0))
    gy = np.where(out, sy * oy / nrm, np.where(qx >= qy, 0.0, sy))
    return phi, gx, gy


def sdf_circle(X, Y, cx, cy, r):
    rx = X - cx
    ry = Y - cy
    d = np.sqrt(rx * rx + ry * ry)
    safe = np.where(d > 0, d, 1.0)
    return d - r, np.where(d > 0, rx / safe, 1.0), np.where(d > 0, ry / safe, 0.0)


def sdf_boom(X, Y, cx, cy, r, hw, length):
    """The oil boom (BoomCase, Main.f90:173-182: r = Dobj/2, hw = Wobj/2,
    length = LBar): a cylinder of radius r with a bar of half width hw
    reaching `length` below where it leaves the cylinder, as the distance to
    their union (the smaller distance, the cylinder's on ties)."""
    p1, x1, y1 = sdf_circle(X, Y, cx, cy, r)
    bottom = cy - np.sqrt(r * r - hw * hw) - length
    p2, x2, y2 = sdf_box(X, Y, cx, 0.5 * (cy + bottom), hw, 0.5 * (cy - bottom))
    bar = p2 < p1
    return np.where(bar, p2, p1), np.where(bar, x2, x1), np.where(bar, y2, y1)


def sdf_ellipse(X, Y, cx, cy, a, b):
    """First-order distance to an ellipse with semi-axes a, b."""
    rx = (X - cx) / a
    ry = (Y - cy) / b
    F = np.sqrt(rx * rx + ry * ry)
    Fs = np.where(F > 0, F, 1.0)
    gx = rx / a / Fs
    gy = ry / b / Fs
    gn = np.sqrt(gx * gx + gy * gy)
    gn = np.where(gn > 0, gn, 1.0)
    return (F - 1.0) / gn, gx / gn, gy / gn


# ---------------------------------------------------------------------------
@dataclass
class Geometry:
    """Cut-cell geometry of the P cells (+ the U/V flags the correction needs)."""
    vofS: np.ndarray
    phiS: np.ndarray
    nxS: np.ndarray
    nyS: np.ndarray
    WlLh: np.ndarray
    aE: np.ndarray   # ghosted (nj+2, ni+2)
    aW: np.ndarray
    aN: np.ndarray
    aS: np.ndarray
    vofSU: np.ndarray
    vofSV: np.ndarray
    moexU: np.ndarray  # int32 cell arrays
    moexV: np.ndarray
    uflag: np.ndarray  # UCell%Posnu/=-1 .and. UCell%MoExCell/=1
    vflag: np.ndarray

    @property
    def fluid(self):
        return self.vofS < 1.0 - EPSI


def _solid_cells(phiS, nxS, nyS, dxc, dyc):
    """Solid volume fraction from the solid plane (Clsvof_Mod.f90:273-277)."""
    from .cutcell import solid_fraction
    return solid_fraction(phiS, nxS, nyS, dxc, dyc)


def solid_planes(g: Grid, solid=None):
    """Solid line of every P, U and V cell: {"P"|"U"|"V": (phiS, nxS, nyS)}
    from ``solid(X, Y) -> (phiS, nxS, nyS)`` (a signed distance, positive in
    the fluid) at the cell centres, or a far plane where there is no solid."""
    from .cutcell import grid_axes
    out = {}
    for k, (x, y, _, _) in grid_axes(g).items():
        X, Y = np.meshgrid(x, y)
        if solid is None:
            out[k] = (np.full(X.shape, 1.0e3), np.zeros(X.shape), np.ones(X.shape))
        else:
            out[k] = tuple(np.ascontiguousarray(a, dtype=np.float64) for a in solid(X, Y))
    return out


def _geometry_from(cc):
    P, U, V = cc["P"], cc["U"], cc["V"]
    uflag = ((U["numbered"] != 0) & (U["moex"] != 1)).astype(np.int32)
    vflag = ((V["numbered"] != 0) & (V["moex"] != 1)).astype(np.int32)
    c = np.ascontiguousarray
    return Geometry(vofS=c(P["vofS"]), phiS=c(P["phiS"]), nxS=c(P["nxS"]), nyS=c(P["nyS"]), WlLh=c(P["WlLh"]),
                    aE=c(P["aE"]), aW=c(P["aW"]), aN=c(P["aN"]), aS=c(P["aS"]), vofSU=c(U["vofS"]),
                    vofSV=c(V["vofS"]), moexU=c(U["moex"].astype(np.int32)), moexV=c(V["moex"].astype(np.int32)),
                    uflag=uflag, vflag=vflag)


def cut_cells(g: Grid, solid=None, prev_AlN=None):
    """The whole cut-cell geometry (Grid_Preprocess + NewCellFace) for a
    solid given as in solid_planes: returns (Geometry, (U, V))."""
    from .cutcell import cut_cell_geometry
    cc = cut_cell_geometry(g, solid_planes(g, solid), prev_AlN=prev_AlN)
    return _geometry_from(cc), tuple(_stag_from(cc[k]) for k in "UV")


def build_geometry(g: Grid, solid=None):
    """Cut-cell geometry of the P cells (+ the U/V flags the correction
    needs) for a solid given as in solid_planes, or no solid at all."""
    return cut_cells(g, solid)[0]


# ---------------------------------------------------------------------------
@dataclass
class Interface:
    vof: np.ndarray   # liquid volume fraction of the whole cell
    phi: np.ndarray   # level set, gas where > 0
    nx: np.ndarray
    ny: np.ndarray
    vofS: np.ndarray = field(default=None)


def interface_from_levelset(g: Grid, geo: Geometry, levelset):
    """vof from the PLIC plane of the level set (liquid where phi < 0),
    scaled by the fluid fraction as Initial_Clsvof does (Clsvof_Mod.f90:393)."""
    Xp, Yp = np.meshgrid(g.xc, g.yc)
    phi, nx, ny = levelset(Xp, Yp)
    DX = np.broadcast_to(g.dx[None, :], g.shape)
    DY = np.broadcast_to(g.dy[:, None], g.shape)
    vof = frac(nx, ny, phi, DX, DY) * (1.0 - geo.vofS)
    return Interface(vof=np.ascontiguousarray(vof), phi=np.ascontiguousarray(phi),
                     nx=np.ascontiguousarray(nx), ny=np.ascontiguousarray(ny), vofS=geo.vofS)


def union_gas(*pieces):
    """Union of gas regions given as (phi, nx, ny) with phi > 0 in gas."""
    phi, nx, ny = pieces[0]
    phi = phi.copy(); nx = nx.copy(); ny = ny.copy()
    for p, a, b in pieces[1:]:
        m = p > phi
        phi = np.where(m, p, phi)
        nx = np.where(m, a, nx)
        ny = np.where(m, b, ny)
    return phi, nx, ny


# ---------------------------------------------------------------------------
# Benchmark configurations (synthetic, seeded).
@dataclass
class Case:
    name: str
    grid: Grid
    fluids: Fluids
    geo: Geometry
    itf: Interface
    dt: float
    tol: float
    bc: tuple = (1, 1, 1, 0)   # W,E,S,N: Neumann W/E/S, Dirichlet N (ProjectionP_Mod.f90:55)
    solid: object = None       # solid(X, Y) -> (phiS, nxS, nyS), or None
    levelset: object = None    # levelset(X, Y) -> (phi, nx, ny), liquid where phi < 0
    shape: tuple = ("none",)   # the solid as (kind, cx, cy, a, b[, c]) for the device geometry rebuild


def _gas_bubble(cx, cy, r):
    def f(X, Y):
        d, gx, gy = sdf_circle(X, Y, cx, cy, r)
        return -d, -gx, -gy  # gas inside the circle
    return f


def case_bubble(ni=128, nj=256):
    """Single bubble r=0.25 at (0.5,0.5) in a 1x2 m liquid column."""
    g = uniform_grid(ni, nj, 1.0, 2.0)
    geo = build_geometry(g, None)
    itf = interface_from_levelset(g, geo, _gas_bubble(0.5, 0.5, 0.25))
    h = g.dx[0]
    return Case("bubble", g, Fluids(), geo, itf, dt=0.25 * h, tol=1e-10, levelset=_gas_bubble(0.5, 0.5, 0.25))


def case_dambreak(ni=512, nj=256):
    """Dam break: water x<1, y<1 in a 4x2 m box, fixed rectangle 0.1x0.3 at (2.5,0.6)."""
    g = uniform_grid(ni, nj, 4.0, 2.0)

    def solid(X, Y):
        return sdf_box(X, Y, 2.5, 0.6, 0.05, 0.15)
    geo = build_geometry(g, solid)

    def ls(X, Y):
        return sdf_box(X, Y, 0.5, 0.5, 0.5, 0.5)  # > 0 outside the water column
    itf = interface_from_levelset(g, geo, ls)
    return Case("dambreak", g, Fluids(), geo, itf, dt=0.25 * g.dx[0], tol=1e-10, solid=solid, levelset=ls,
                shape=("box", 2.5, 0.6, 0.05, 0.15))


def swarm_bubbles(seed=0, n=64, depth=0.8, rmin=0.01, rmax=0.03, ymin=0.1, ymax=0.6):
    rng = np.random.default_rng(seed)
    out = []
    tries = 0
    while len(out) < n and tries < 100000:
        tries += 1
        r = rng.uniform(rmin, rmax)
        cy = rng.uniform(ymin, ymax)
        cx = rng.uniform(r, 1.0 - r)
        if all((cx - a) ** 2 + (cy - b) ** 2 > (r + c) ** 2 for a, b, c in out):
            out.append((cx, cy, r))
    return out


def case_swarm(ni=2048, nj=2048, seed=0):
    """64 seeded bubbles in a 1x1 m box filled to 0.8 m."""
    g = uniform_grid(ni, nj, 1.0, 1.0)
    geo = build_geometry(g, None)
    bubbles = swarm_bubbles(seed)

    def ls(X, Y):
        # union of the gas above the surface and the bubbles; each bubble is
        # evaluated in a window of a few cells around it, which keeps the
        # sign everywhere and the distance exact in the interface band
        phi = Y - 0.8
        nx = np.zeros(X.shape)
        ny = np.ones(X.shape)
        h = max(float(g.dx.max()), float(g.dy.max()))
        x1, y1 = X[0], Y[:, 0]
        for cx, cy, r in bubbles:
            m = r + max(8 * h, 0.05)
            i0, i1 = np.searchsorted(x1, cx - m), np.searchsorted(x1, cx + m)
            j0, j1 = np.searchsorted(y1, cy - m), np.searchsorted(y1, cy + m)
            sl = (slice(j0, j1), slice(i0, i1))
            p, a, b = _gas_bubble(cx, cy, r)(X[sl], Y[sl])
            sel = p > phi[sl]
            phi[sl] = np.where(sel, p, phi[sl])
            nx[sl] = np.where(sel, a, nx[sl])
            ny[sl] = np.where(sel, b, ny[sl])
        return phi, nx, ny
    itf = interface_from_levelset(g, geo, ls)
    return Case("swarm", g, Fluids(), geo, itf, dt=0.25 * g.dx[0], tol=1e-8, levelset=ls)


def case_waves(ni=1024, nj=1024, seed=0):
    """8x4 m tank, surface y = 2 + 8 seeded sinusoids, ellipse 0.4x0.15 at the surface."""
    g = uniform_grid(ni, nj, 8.0, 4.0)
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.01, 0.1, 8)
    lam = rng.uniform(0.5, 4.0, 8)
    th = rng.uniform(0.0, 2 * np.pi, 8)

    def surf(x):
        return 2.0 + np.sum(a[:, None] * np.sin(2 * np.pi * x[None, :] / lam[:, None] + th[:, None]), axis=0)

    def dsurf(x):
        return np.sum(a[:, None] * (2 * np.pi / lam[:, None]) *
                      np.cos(2 * np.pi * x[None, :] / lam[:, None] + th[:, None]), axis=0)
    yc = float(surf(np.array([4.0]))[0])

    def solid(X, Y):
        return sdf_ellipse(X, Y, 4.0, yc, 0.2, 0.075)
    geo = build_geometry(g, solid)

    def ls(X, Y):
        x = X[0]
        s = surf(x)[None, :]
        ds = dsurf(x)[None, :]
        nrm = np.sqrt(1.0 + ds * ds)
        return (Y - s) / nrm, np.broadcast_to(-ds / nrm, X.shape).copy(), np.broadcast_to(1.0 / nrm, X.shape).copy()
    itf = interface_from_levelset(g, geo, ls)
    return Case("waves", g, Fluids(), geo, itf, dt=0.25 * g.dx[0], tol=1e-8, solid=solid, levelset=ls,
                shape=("ellipse", 4.0, yc, 0.2, 0.075))


CASES = {"bubble": case_bubble, "dambreak": case_dambreak, "swarm": case_swarm, "waves": case_waves}


# ---------------------------------------------------------------------------
def predictor_fields(case: Case, seed=1, amp=0.1):
    """Synthetic predictor outputs for a projection: Pred%u/v (ghosted),
    PU%Dp / PV%Dp (ghosted).  The velocities are a smooth seeded field plus
    the gravity increment -g*dt, with walls closed (u(0,j)=u(ni,j)=0,
    v(i,0)=0); Dp is dt scaled by a smooth positive factor, the form
    dx*dy/(aP - sum a_nb) of PredictorUV_mod.f90:700 takes with viscous and
    wall terms."""
    g = case.grid
    rng = np.random.default_rng(seed)
    nj, ni = g.nj, g.ni
    X, Y = np.meshgrid(np.arange(ni + 2) / max(ni, 1), np.arange(nj + 2) / max(nj, 1))
    k = rng.uniform(1.0, 4.0, size=4)
    ph = rng.uniform(0, 2 * np.pi, size=4)
    u = amp * np.sin(2 * np.pi * k[0] * X + ph[0]) * np.cos(2 * np.pi * k[1] * Y + ph[1])
    v = amp * np.cos(2 * np.pi * k[2] * X + ph[2]) * np.sin(2 * np.pi * k[3] * Y + ph[3])
    v = v - case.fluids.g * case.dt
    u[:, 0] = 0.0
    u[:, ni] = 0.0
    u[:, ni + 1] = 0.0
    v[0, :] = 0.0
    DpU = case.dt / (1.0 + 0.05 * (1 + np.sin(2 * np.pi * (X + 0.3 * Y))))
    DpV = case.dt / (1.0 + 0.05 * (1 + np.cos(2 * np.pi * (0.7 * X - Y))))
    return (np.ascontiguousarray(u), np.ascontiguousarray(v),
            np.ascontiguousarray(DpU), np.ascontiguousarray(DpV))


# ---------------------------------------------------------------------------
def transport_velocity(g: Grid, kind="rotation", amp=1.0, seed=0, walls=True):
    """Face velocities TVar%u / TVar%v (ghosted, (nj+2, ni+2)) for the
    interface transport.  u(i,j) sits on the east face of P cell (i,j),
    v(i,j) on its north face (Mesh_Mod.f90:151,165).

    kind: "rotation" solid-body rotation about the domain centre;
          "uniform"  constant (amp, amp/2);
          "vortex"   seeded sum of modes of a node stream function, so the
                     discrete divergence (u_e-u_w)/dx + (v_n-v_s)/dy vanishes;
          "random"   seeded white noise (not divergence-free).
    walls: zero normal velocity on the four boundary faces."""
    ni, nj = g.ni, g.nj
    xf = g.xc[0] - g.dx[0] / 2 + np.concatenate([[0.0], np.cumsum(g.dx)])  # faces x_{i+1/2}, i=0..ni
    yf = g.yc[0] - g.dy[0] / 2 + np.concatenate([[0.0], np.cumsum(g.dy)])
    u = np.zeros((nj + 2, ni + 2))
    v = np.zeros((nj + 2, ni + 2))
    xm, ym = 0.5 * (xf[0] + xf[-1]), 0.5 * (yf[0] + yf[-1])
    if kind == "rotation":
        Yc, Xf = np.meshgrid(g.yc, xf, indexing="ij")       # (nj, ni+1)
        u[1:nj + 1, 0:ni + 1] = -amp * (Yc - ym)
        Yf, Xc = np.meshgrid(yf, g.xc, indexing="ij")       # (nj+1, ni)
        v[0:nj + 1, 1:ni + 1] = amp * (Xc - xm)
    elif kind == "uniform":
        u[1:nj + 1, 0:ni + 1] = amp
        v[0:nj + 1, 1:ni + 1] = 0.5 * amp
    elif kind == "vortex":
        rng = np.random.default_rng(seed)
        Lx, Ly = xf[-1] - xf[0], yf[-1] - yf[0]
        Yn, Xn = np.meshgrid(yf, xf, indexing="ij")         # nodes (nj+1, ni+1)
        psi = np.zeros(Xn.shape)
        for _ in range(4):
            kx, ky = rng.integers(1, 4, size=2)
            a = rng.uniform(-1, 1)
            psi += a * np.sin(kx * np.pi * (Xn - xf[0]) / Lx) * np.sin(ky * np.pi * (Yn - yf[0]) / Ly)
        psi *= amp * min(Lx, Ly) / np.pi
        u[1:nj + 1, 0:ni + 1] = (psi[1:, :] - psi[:-1, :]) / g.dy[:, None]
        v[0:nj + 1, 1:ni + 1] = -(psi[:, 1:] - psi[:, :-1]) / g.dx[None, :]
    elif kind == "random":
        rng = np.random.default_rng(seed)
        u[:, :] = amp * rng.uniform(-1, 1, u.shape)
        v[:, :] = amp * rng.uniform(-1, 1, v.shape)
    else:
        raise ValueError(kind)
    if walls:
        u[:, 0] = 0.0
        u[:, ni] = 0.0
        v[0, :] = 0.0
        v[nj, :] = 0.0
    return np.ascontiguousarray(u), np.ascontiguousarray(v)


# ---------------------------------------------------------------------------
# Staggered (U / V) cell geometry for the momentum predictor.
@dataclass
class StagGeometry:
    """Cut-cell geometry of one staggered grid, in the arrays the predictor
    reads (PredictorUV_mod.f90).  Cell arrays (nj, ni): vofS, nxS, nyS, phiS,
    WlLh, delh, ccx, ccy (fluid centroid offset from the cell centre), moex,
    active (a row of the momentum system).  Ghosted arrays (nj+2, ni+2):
    apertures aE..aS, EtaE, EtaN, AlE, AlN, DAlE, DAlN, SxE, SyN and the
    open-face centres FCE (fcex, fcey) and FCN (fcnx, fcny), offsets from
    the cell centre."""
    vofS: np.ndarray
    nxS: np.ndarray
    nyS: np.ndarray
    phiS: np.ndarray
    WlLh: np.ndarray
    delh: np.ndarray
    ccx: np.ndarray
    ccy: np.ndarray
    moex: np.ndarray
    active: np.ndarray
    aE: np.ndarray
    aW: np.ndarray
    aN: np.ndarray
    aS: np.ndarray
    EtaE: np.ndarray
    EtaN: np.ndarray
    AlE: np.ndarray
    AlN: np.ndarray
    DAlE: np.ndarray
    DAlN: np.ndarray
    SxE: np.ndarray
    SyN: np.ndarray
    fcex: np.ndarray
    fcey: np.ndarray
    fcnx: np.ndarray
    fcny: np.ndarray


def _stag_from(T):
    c = np.ascontiguousarray
    kw = {k: c(T[k]) for k in ("vofS", "nxS", "nyS", "phiS", "WlLh", "delh", "ccx", "ccy", "aE", "aW", "aN", "aS",
                               "EtaE", "EtaN", "AlE", "AlN", "DAlE", "DAlN", "SxE", "SyN", "fcex", "fcey", "fcnx",
                               "fcny")}
    kw["moex"] = c(T["moex"].astype(np.int32))
    kw["active"] = c(T["active"].astype(np.int32))
    return StagGeometry(**kw)


def staggered_geometry(g: Grid, geo: Geometry = None, solid=None, prev_AlN=None):
    """U and V cell geometry for the predictor (NewCellFace); `solid` as in
    solid_planes.  prev_AlN = {"U": AlN, "V": AlN} of the geometry before
    (AlE(0,j) reads it); by default the same geometry's own."""
    return cut_cells(g, solid, prev_AlN)[1]


def staggered_vof(g: Grid, stag: StagGeometry, which, levelset):
    """Liquid fraction of U (which=0) or V cells from a level set (liquid
    where phi < 0), scaled by the fluid fraction."""
    ni, nj = g.ni, g.nj
    DX = np.broadcast_to(g.dx[None, :], (nj, ni))
    DY = np.broadcast_to(g.dy[:, None], (nj, ni))
    Xp, Yp = np.meshgrid(g.xc, g.yc)
    if which == 0:
        X, Y, hx, hy = Xp + 0.5 * DX, Yp, np.broadcast_to(g.udx[None, :], (nj, ni)), DY
    else:
        X, Y, hx, hy = Xp, Yp + 0.5 * DY, DX, np.broadcast_to(g.vdy[:, None], (nj, ni))
    phi, nx, ny = levelset(X, Y)
    return np.ascontiguousarray(frac(nx, ny, phi, hx, hy) * (1.0 - stag.vofS))
