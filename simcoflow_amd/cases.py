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
import heapq

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
    vofS = frac(nxS, nyS, phiS, dxc, dyc)
    vofS = np.where(vofS < VOFEPS, 0.0, vofS)
    vofS = np.where(vofS >= 1.0 - VOFEPS, 1.0, vofS)
    return vofS


def _edge_areas(phiS, nxS, nyS, dxc, dyc):
    """Edge_Geo_Cal on the four edges (Cutcell_Mod.f90:52-73,302-337):
    returns the *solid* fraction of each edge (W, N, E, S)."""
    hx = 0.5 * dxc
    hy = 0.5 * dyc
    eps = 1e-14
    # corner level-set values: SW, NW, NE, SE
    L = {}
    for key, (sx, sy) in {"sw": (-1, -1), "nw": (-1, 1), "ne": (1, 1), "se": (1, -1)}.items():
        L[key] = nxS * (sx * hx) + nyS * (sy * hy) + phiS

    def edge(l1, l2):
        a = np.zeros(np.broadcast(l1, l2).shape)
        both_solid = (l1 < eps) & (l2 < eps)
        a[both_solid] = 1.0
        m1 = (l1 >= eps) & (l2 < eps)
        m2 = (l1 < eps) & (l2 >= eps)
        den = np.abs(l1) + np.abs(l2)
        den = np.where(den > 0, den, 1.0)
        a = np.where(m1, np.abs(l2) / den, a)
        a = np.where(m2, np.abs(l1) / den, a)
        return a

    return (edge(L["sw"], L["nw"]), edge(L["nw"], L["ne"]),
            edge(L["ne"], L["se"]), edge(L["se"], L["sw"]))


def _cell_geo(phiS, nxS, nyS, vofS, dxc, dyc, closed_rule):
    """Aperture part of Cell_Geo_Cal (Cutcell_Mod.f90:34-300) for one grid.
    Returns ghosted apertures, the wall length and the modified vofS."""
    nj, ni = phiS.shape
    sW, sN, sE, sS = _edge_areas(phiS, nxS, nyS, dxc, dyc)
    # wall length from the raw edge values (Cutcell_Mod.f90:79-89)
    WlLh = np.sqrt(((sE - sW) * dxc) ** 2 + ((sN - sS) * dyc) ** 2)
    aW = np.zeros((nj + 2, ni + 2)); aE = np.zeros_like(aW)
    aN = np.zeros_like(aW); aS = np.zeros_like(aW)
    aW[1:-1, 1:-1] = 1.0 - sW
    aE[1:-1, 1:-1] = 1.0 - sE
    aN[1:-1, 1:-1] = 1.0 - sN
    aS[1:-1, 1:-1] = 1.0 - sS
    # harmonic face averaging for i=2..Isize-1, j=2..Jsize-1 (:161-204)
    if ni >= 3 and nj >= 3:
        I = slice(2, ni)      # ghosted index of i = 2..ni-1
        J = slice(2, nj)
        Ip = slice(3, ni + 1)
        Jp = slice(3, nj + 1)
        n_ = aN[J, I]; s_ = aS[Jp, I]
        avg = 2.0 * n_ * s_ / (n_ + s_ + ROUNDOFF)
        aN[J, I] = avg
        aS[Jp, I] = avg
        e_ = aE[J, I]; w_ = aW[J, Ip]
        avg = 2.0 * e_ * w_ / (e_ + w_ + ROUNDOFF)
        aE[J, I] = avg
        aW[J, Ip] = avg
    vofS = vofS.copy()
    WlLh = WlLh.copy()
    # solid cells close their faces and their neighbours' (:215-228)
    solid = vofS > 1.0 - EPSI
    vofS[solid] = 1.0
    js, is_ = np.nonzero(solid)
    gj, gi = js + 1, is_ + 1
    for arr in (aW, aE, aN, aS):
        arr[gj, gi] = 0.0
    aE[gj, gi - 1] = 0.0
    aN[gj - 1, gi] = 0.0
    aW[gj, gi + 1] = 0.0
    aS[gj + 1, gi] = 0.0
    if closed_rule:
        # cells closed on both sides of an axis become solid, visited in the
        # reference's loop order (i outer, j inner; :229-242)
        def closed(j, i):
            g_j, g_i = j + 1, i + 1
            return ((aW[g_j, g_i] + aE[g_j, g_i] < ROUNDOFF or aS[g_j, g_i] + aN[g_j, g_i] < ROUNDOFF)
                    and vofS[j, i] < 1.0 - EPSI)
        anyzero = ((aW[1:-1, 1:-1] == 0) | (aE[1:-1, 1:-1] == 0) | (aN[1:-1, 1:-1] == 0) |
                   (aS[1:-1, 1:-1] == 0)) & (vofS < 1.0 - EPSI)
        cj, ci = np.nonzero(anyzero)
        heap = [(int(i), int(j)) for j, i in zip(cj, ci)]
        heapq.heapify(heap)
        seen = set(heap)
        while heap:
            i, j = heapq.heappop(heap)
            if closed(j, i):
                vofS[j, i] = 1.0
                g_j, g_i = j + 1, i + 1
                for arr in (aW, aE, aN, aS):
                    arr[g_j, g_i] = 0.0
                aE[g_j, g_i - 1] = 0.0
                aN[g_j - 1, g_i] = 0.0
                aW[g_j, g_i + 1] = 0.0
                aS[g_j + 1, g_i] = 0.0
                for (ii, jj) in ((i + 1, j), (i, j + 1)):
                    if ii < ni and jj < nj and (ii, jj) not in seen:
                        seen.add((ii, jj))
                        heapq.heappush(heap, (ii, jj))
    # wall length of solid cells (:245-258)
    full = np.abs(vofS - 1.0) < ROUNDOFF
    ew0 = (aW[1:-1, 1:-1] == 0.0) | (aE[1:-1, 1:-1] == 0.0)
    ns0 = (aN[1:-1, 1:-1] == 0.0) | (aS[1:-1, 1:-1] == 0.0)
    WlLh = np.where(full & ew0, np.broadcast_to(dyc, WlLh.shape), WlLh)
    WlLh = np.where(full & ns0, np.broadcast_to(dxc, WlLh.shape), WlLh)
    return aE, aW, aN, aS, WlLh, vofS


def build_geometry(g: Grid, solid=None):
    """Cut-cell geometry for a solid given as ``solid(X, Y) -> (phiS, nxS, nyS)``
    (a signed distance, positive in the fluid), or no solid at all."""
    ni, nj = g.ni, g.nj
    DX = np.broadcast_to(g.dx[None, :], (nj, ni))
    DY = np.broadcast_to(g.dy[:, None], (nj, ni))
    Xp, Yp = np.meshgrid(g.xc, g.yc)
    # U cells centred on east faces, V cells on north faces (Mesh_Mod.f90:151,165)
    Xu = Xp + 0.5 * DX
    Yv = Yp + 0.5 * DY
    UDX = np.broadcast_to(g.udx[None, :], (nj, ni))
    VDY = np.broadcast_to(g.vdy[:, None], (nj, ni))

    if solid is None:
        # no solid: every aperture open (Edge_Geo_Cal gives 0 solid on every
        # edge), no wall, no momentum-exchange cell
        zc = np.zeros((nj, ni))
        ones_g = np.zeros((nj + 2, ni + 2))
        ones_g[1:-1, 1:-1] = 1.0
        return Geometry(vofS=zc.copy(), phiS=np.full((nj, ni), 1.0e3), nxS=zc.copy(), nyS=np.ones((nj, ni)),
                        WlLh=zc.copy(), aE=ones_g.copy(), aW=ones_g.copy(), aN=ones_g.copy(), aS=ones_g.copy(),
                        vofSU=zc.copy(), vofSV=zc.copy(), moexU=np.zeros((nj, ni), np.int32),
                        moexV=np.zeros((nj, ni), np.int32), uflag=np.ones((nj, ni), np.int32),
                        vflag=np.ones((nj, ni), np.int32))

    def plane(X, Y):
        return solid(X, Y)

    phiS, nxS, nyS = plane(Xp, Yp)
    vofS = _solid_cells(phiS, nxS, nyS, DX, DY)
    aE, aW, aN, aS, WlLh, vofS = _cell_geo(phiS, nxS, nyS, vofS, DX, DY, closed_rule=True)

    phiU, nxU, nyU = plane(Xu, Yp)
    vofSU = _solid_cells(phiU, nxU, nyU, UDX, DY)
    _, _, _, _, _, vofSU = _cell_geo(phiU, nxU, nyU, vofSU, UDX, DY, closed_rule=False)
    phiV, nxV, nyV = plane(Xp, Yv)
    vofSV = _solid_cells(phiV, nxV, nyV, DX, VDY)
    _, _, _, _, _, vofSV = _cell_geo(phiV, nxV, nyV, vofSV, DX, VDY, closed_rule=False)

    # DefineMomentumExchangeCell (Cutcell_Mod.f90:360-397), in loop order
    moexU = np.zeros((nj, ni), np.int32)
    moexV = np.zeros((nj, ni), np.int32)
    candU = (vofSU >= EPSI) & (vofSU < 1.0 - EPSI)
    candV = (vofSV >= EPSI) & (vofSV < 1.0 - EPSI)
    cj, ci = np.nonzero(candU | candV)
    order = np.lexsort((cj, ci))
    for k in order:
        j, i = int(cj[k]), int(ci[k])
        gj, gi = j + 1, i + 1
        if candU[j, i]:
            if i < ni - 1:
                if vofS[j, i] >= 1 - EPSI or vofS[j, i + 1] >= 1 - EPSI or aE[gj, gi] < EPSIF:
                    moexU[j, i] = 1
                    aE[gj, gi] = 0.0
                    aW[gj, gi + 1] = 0.0
            elif vofS[j, i] >= 1 - EPSI:
                moexU[j, i] = 1
        if candV[j, i]:
            if j < nj - 1:
                if vofS[j, i] >= 1 - EPSI or vofS[j + 1, i] >= 1 - EPSI or aN[gj, gi] < EPSIF:
                    moexV[j, i] = 1
                    aN[gj, gi] = 0.0
                    aS[gj + 1, gi] = 0.0
            elif vofS[j, i] >= 1 - EPSI:
                moexV[j, i] = 1
    uflag = ((vofSU < 1.0 - EPSI) & (moexU != 1)).astype(np.int32)
    vflag = ((vofSV < 1.0 - EPSI) & (moexV != 1)).astype(np.int32)
    return Geometry(vofS=np.ascontiguousarray(vofS), phiS=np.ascontiguousarray(phiS),
                    nxS=np.ascontiguousarray(nxS), nyS=np.ascontiguousarray(nyS),
                    WlLh=np.ascontiguousarray(WlLh), aE=aE, aW=aW, aN=aN, aS=aS,
                    vofSU=np.ascontiguousarray(vofSU), vofSV=np.ascontiguousarray(vofSV),
                    moexU=moexU, moexV=moexV, uflag=uflag, vflag=vflag)


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


def _fluid_centroid(nxS, nyS, phiS, hx, hy):
    """Centroid of {nxS x + nyS y + phiS >= 0} inside the hx*hy box centred
    at 0 (polygon clipping); (0,0) when the box is all fluid or all solid."""
    poly = [(-hx / 2, -hy / 2), (hx / 2, -hy / 2), (hx / 2, hy / 2), (-hx / 2, hy / 2)]
    out = []
    for k in range(4):
        p, q = poly[k], poly[(k + 1) % 4]
        fp = nxS * p[0] + nyS * p[1] + phiS
        fq = nxS * q[0] + nyS * q[1] + phiS
        if fp >= 0:
            out.append(p)
        if (fp < 0 <= fq) or (fq < 0 <= fp):
            t = fp / (fp - fq)
            out.append((p[0] + t * (q[0] - p[0]), p[1] + t * (q[1] - p[1])))
    if len(out) < 3:
        return 0.0, 0.0
    a = cx = cy = 0.0
    for k in range(len(out)):
        p, q = out[k], out[(k + 1) % len(out)]
        cr = p[0] * q[1] - q[0] * p[1]
        a += cr
        cx += (p[0] + q[0]) * cr
        cy += (p[1] + q[1]) * cr
    if abs(a) < 1e-300:
        return 0.0, 0.0
    return cx / (3.0 * a), cy / (3.0 * a)


def _stag(g, hx, hy, phiS, nxS, nyS, vofS, moex, which, spacing_next):
    """One staggered grid.  hx, hy: (nj, ni) cell sizes; spacing_next: the
    centre distance to the next cell along x (for SxE) and along y (SyN)."""
    nj, ni = phiS.shape
    aE, aW, aN, aS, WlLh, vofS2 = _cell_geo(phiS, nxS, nyS, vofS, hx, hy, closed_rule=False)
    cut = (vofS2 > EPSI) & (vofS2 < 1.0 - EPSI)
    ccx = np.zeros((nj, ni))
    ccy = np.zeros((nj, ni))
    for j, i in zip(*np.nonzero(cut)):
        ccx[j, i], ccy[j, i] = _fluid_centroid(nxS[j, i], nyS[j, i], phiS[j, i], hx[j, i], hy[j, i])
    G = lambda: np.zeros((nj + 2, ni + 2))
    fcex, fcey, fcnx, fcny = G(), G(), G(), G()
    fcex[1:-1, 1:-1] = hx / 2
    fcny[1:-1, 1:-1] = hy / 2
    # open segment of a cut east / north edge: its centre sits at half the open
    # length from the fluid end
    e_open = aE[1:-1, 1:-1]
    n_open = aN[1:-1, 1:-1]
    fcey[1:-1, 1:-1] = np.where(cut, (0.5 - e_open / 2.0) * np.sign(nyS) * hy, 0.0)
    fcnx[1:-1, 1:-1] = np.where(cut, (0.5 - n_open / 2.0) * np.sign(nxS) * hx, 0.0)
    for a in (fcex, fcey, fcnx, fcny):
        a[0, :] = a[1, :]
        a[-1, :] = a[-2, :]
        a[:, 0] = a[:, 1]
        a[:, -1] = a[:, -2]
    active = (vofS2 < 1.0 - EPSI).astype(np.int32)
    if which == 0:
        active[:, -1] = 0
    else:
        active[-1, :] = 0
    SxE, EtaE, AlE, DAlE = G(), G(), G(), G()
    SyN, EtaN, AlN, DAlN = G(), G(), G(), G()
    sx_next, sy_next = spacing_next
    SxE[1:-1, 1:-1] = sx_next
    SyN[1:-1, 1:-1] = sy_next
    EtaE[:] = 0.5
    EtaN[:] = 0.5
    AlE[:] = 1.0
    AlN[:] = 1.0
    # cut faces: interpolation factor from the centroids, the shift of the
    # interpolation point to the face centre (EastFaceInterpolationInf,
    # NorthFaceInterpolationInf of Cutcell_Mod.f90 in simplified form)
    for j in range(nj):
        for i in range(ni):
            if i + 1 < ni and active[j, i] and active[j, i + 1] and (cut[j, i] or cut[j, i + 1]):
                sx = ccx[j, i + 1] + sx_next[j, i] - ccx[j, i]
                sy = ccy[j, i + 1] - ccy[j, i]
                SxE[j + 1, i + 1] = sx
                eta = abs(fcex[j + 1, i + 1] - ccx[j, i]) / sx
                EtaE[j + 1, i + 1] = 0.5 if eta >= 1.0 else eta
                if abs(sy) > 1e-4 * hy[j, i]:
                    AlE[j + 1, i + 1] = min(2.0, 1.0 + abs(sy) / max(hy[j, i], 1e-300))
                    DAlE[j + 1, i + 1] = sy * 0.5 * (nyS[j, i] + nyS[j, i + 1]) / (sx * hy[j, i])
            if j + 1 < nj and active[j, i] and active[j + 1, i] and (cut[j, i] or cut[j + 1, i]):
                sy = ccy[j + 1, i] + sy_next[j, i] - ccy[j, i]
                sx = ccx[j + 1, i] - ccx[j, i]
                SyN[j + 1, i + 1] = sy
                eta = abs(fcny[j + 1, i + 1] - ccy[j, i]) / sy
                EtaN[j + 1, i + 1] = 0.5 if eta >= 1.0 else eta
                if abs(sx) >= 1e-4 * hx[j, i]:
                    AlN[j + 1, i + 1] = min(2.0, 1.0 + abs(sx) / max(hx[j, i], 1e-300))
                    DAlN[j + 1, i + 1] = sx * 0.5 * (nxS[j, i] + nxS[j + 1, i]) / (sy * hx[j, i])
    for a in (SxE, EtaE, AlE, DAlE, SyN, EtaN, AlN, DAlN):
        a[0, :] = a[1, :]
        a[-1, :] = a[-2, :]
        a[:, 0] = a[:, 1]
        a[:, -1] = a[:, -2]
    delh = np.abs(ccx * nxS + ccy * nyS + phiS) + 1.0e-24
    delh = np.where(moex == 1, 1.0e-20, delh)
    c = np.ascontiguousarray
    return StagGeometry(vofS=c(vofS2), nxS=c(nxS), nyS=c(nyS), phiS=c(phiS), WlLh=c(WlLh), delh=c(delh),
                        ccx=c(ccx), ccy=c(ccy), moex=c(moex.astype(np.int32)), active=c(active),
                        aE=aE, aW=aW, aN=aN, aS=aS, EtaE=EtaE, EtaN=EtaN, AlE=AlE, AlN=AlN, DAlE=DAlE,
                        DAlN=DAlN, SxE=SxE, SyN=SyN, fcex=fcex, fcey=fcey, fcnx=fcnx, fcny=fcny)


def staggered_geometry(g: Grid, geo: Geometry, solid=None):
    """U and V cell geometry for the predictor; `solid` as in build_geometry."""
    ni, nj = g.ni, g.nj
    DX = np.broadcast_to(g.dx[None, :], (nj, ni))
    DY = np.broadcast_to(g.dy[:, None], (nj, ni))
    UDX = np.ascontiguousarray(np.broadcast_to(g.udx[None, :], (nj, ni)))
    VDY = np.ascontiguousarray(np.broadcast_to(g.vdy[:, None], (nj, ni)))
    Xp, Yp = np.meshgrid(g.xc, g.yc)
    Xu, Yv = Xp + 0.5 * DX, Yp + 0.5 * DY
    if solid is None:
        zero = np.zeros((nj, ni))
        planes = [(np.full((nj, ni), 1.0e3), zero.copy(), np.ones((nj, ni)))] * 2
    else:
        planes = [solid(Xu, Yp), solid(Xp, Yv)]
    # centre distance to the next cell: U cells along x = the P spacing of
    # cell i+1, along y = dy; V cells along x = the U spacing, along y = dy
    # of cell j+1 (Cutcell_Mod.f90 EastFaceInterpolationInf BGrid%dx)
    dxn = np.append(g.dx[1:], g.dx[-1])[None, :].repeat(nj, 0)
    dyn = np.append(g.dy[1:], g.dy[-1])[:, None].repeat(ni, 1)
    U = _stag(g, UDX, np.ascontiguousarray(DY), planes[0][0], planes[0][1], planes[0][2], geo.vofSU.copy(),
              geo.moexU, 0, (dxn, np.ascontiguousarray(VDY)))
    V = _stag(g, np.ascontiguousarray(DX), VDY, planes[1][0], planes[1][1], planes[1][2], geo.vofSV.copy(),
              geo.moexV, 1, (UDX, dyn))
    return U, V


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
