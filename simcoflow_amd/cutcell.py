"""Cut-cell geometry of the P, U and V cells on the host (numpy).

The Python host's form of the reference's ``Grid_Preprocess`` followed by
``NewCellFace`` (the order of ``Solver_Mod.f90:303-304``), evaluated array
by array with every expression in the reference's order, so the result is
the same, bit for bit, as a cell-by-cell evaluation:

* ``Edge_Geo_Cal`` (``Cutcell_Mod.f90:302-337``) on the four edges of every
  cell: solid fraction of the edge and the centre of its solid part;
* ``Cell_Geo_Cal`` (``:34-300``): wall length, face-centre ghosts, fluid
  fractions and fluid face centres (``:119-159``), harmonic face averaging
  of ``i = 2..Isize-1, j = 2..Jsize-1`` (``:161-204``), solid cells closing
  their faces and, on the P cells, cells shut on both sides of an axis
  becoming solid (``:214-243``), the wall length of full solid cells
  (``:245-258``) and the corner patch under flat solid bottoms (ICorProb,
  ``:259-286``);
* ``DefineMomentumExchangeCell`` (``:360-397``) and ``NumberExternalCell``
  (``:340-358``);
* ``NewCellFace`` (``:399-549``): the U/V centroids from the P cells' face
  centres, the small cells next to one P cell, the wall distance ``delh``
  and the wall length from the final apertures, then
  ``EastFaceInterpolationInf`` / ``NorthFaceInterpolationInf``
  (``:551-700``).

Two steps of the reference are sequential in loop order and are recast:

* the closing loop visits cells with i outer and j inner; when (i, j) is
  visited its west and south faces carry the closings of (i-1, j) and
  (i, j-1), its east and north faces are still the averaged values.  The
  decisions therefore solve a recurrence over an acyclic graph whose update
  is monotone, and iterating it from the solid cells reaches that solution;
* ``AlE(0,j) = AlN(1,j)`` (``:606``) reads the north factor before the same
  call's ``NorthFaceInterpolationInf`` rewrites it, i.e. the previous call's
  value: ``prev_AlN`` passes it; without it the same geometry's own value is
  used (what every call after the first reads when the solid is at rest).

Storage the reference reads without defining it gets a fixed value: aperture
ghosts 0; Posnu of the U cells at i = Isize and of the V cells at j = Jsize
(``NumberExternalCell`` stops one short) the 0 of a fresh allocation, so
those cells count as numbered wherever Posnu is tested (they have no row in
the momentum systems); the face-factor ghosts neither interpolation routine
writes take the nearest written entry, rows first.

Arrays use the reference's Fortran layout: cell arrays (nj, ni), ghosted
arrays (nj+2, ni+2) with the (0,0) ghost at [0, 0].
"""
import numpy as np

EPSI = 1.0e-3       # StateVariables.f90:28
EPSIF = 1.0e-2      # StateVariables.f90:28
VOFEPS = 1.0e-14    # Clsvof_Mod.f90:10
ROUNDOFF = 1.0e-24  # Cutcell_Mod.f90:45
TOL = 1.0e-24       # Cutcell_Mod.f90:410


def grid_axes(g):
    """(x, y, dx, dy) per i / per j of the P, U and V grids
    (Mesh_Mod.f90:123-175): U cells centred on the east faces of the P
    cells, V cells on their north faces."""
    return {"P": (g.xc, g.yc, g.dx, g.dy),
            "U": (g.xc + 0.5 * g.dx, g.yc, g.udx, g.dy),
            "V": (g.xc, g.yc + 0.5 * g.dy, g.dx, g.vdy)}


def _frac(nx, ny, d, dx, dy):
    from .cases import frac
    return frac(nx, ny, d, dx, dy)


def solid_fraction(phiS, nxS, nyS, DX, DY):
    """frac of the solid line, then the clipping of Clsvof_Mod.f90:273-277."""
    v = _frac(nxS, nyS, phiS, DX, DY)
    v = np.where(v < VOFEPS, 0.0, v)
    return np.where(v >= 1.0 - VOFEPS, 1.0, v)


def edge_geo(p1x, p1y, p2x, p2y, nx, ny, d):
    """Edge_Geo_Cal (Cutcell_Mod.f90:302-337) on arrays: solid fraction of
    the edge p1-p2 and the centre (fx, fy) of its solid part."""
    l1 = nx * p1x + ny * p1y + d
    l2 = nx * p2x + ny * p2y + d
    eps = 1.0e-14
    f1 = l1 >= eps
    f2 = l2 >= eps
    s1 = l1 < eps
    s2 = l2 < eps
    fx = np.broadcast_to(0.5 * (p1x + p2x), l1.shape).copy()
    fy = np.broadcast_to(0.5 * (p1y + p2y), l1.shape).copy()
    a = np.zeros(l1.shape)
    a[s1 & s2] = 1.0
    m1 = f1 & s2     # Pt2 in the solid
    m2 = s1 & f2     # Pt1 in the solid
    with np.errstate(invalid="ignore", divide="ignore"):
        den = np.abs(l1) + np.abs(l2)
        a1 = np.abs(l2) / den
        a2 = np.abs(l1) / den
    a = np.where(m1, a1, np.where(m2, a2, a))
    vert = np.broadcast_to(np.abs(p1x - p2x) <= 1.0e-10, l1.shape)
    P1x, P1y, P2x, P2y = (np.broadcast_to(p, l1.shape) for p in (p1x, p1y, p2x, p2y))
    fy = np.where(m1 & vert, P2y - 0.5 * a * (P2y - P1y), fy)
    fx = np.where(m1 & ~vert, P2x - 0.5 * a * (P2x - P1x), fx)
    fy = np.where(m2 & vert, P1y - 0.5 * a * (P1y - P2y), fy)
    fx = np.where(m2 & ~vert, P1x - 0.5 * a * (P1x - P2x), fx)
    return a, fx, fy


def _ghosted(nj, ni):
    return np.zeros((nj + 2, ni + 2))


def cell_geo_cal(axes, phiS, nxS, nyS, vofS, p_cell, icorprob=True):
    """Cell_Geo_Cal (Cutcell_Mod.f90:34-300) for one grid.  Returns a dict
    with the ghosted apertures aE/aW/aN/aS and face centres fcex/fcey/fcnx/
    fcny, the wall length WlLh, the closed vofS, and (P cells, corner patch)
    nxS/nyS."""
    x, y, dx, dy = axes
    nj, ni = phiS.shape
    DX = np.broadcast_to(dx[None, :], (nj, ni))
    DY = np.broadcast_to(dy[:, None], (nj, ni))
    X0 = -(DX * 0.5)
    X1 = -(DX * -0.5)
    Y0 = -(DY * 0.5)
    Y1 = -(DY * -0.5)
    # W: Pt(0,0)-Pt(0,1); N: Pt(0,1)-Pt(1,1); E: Pt(1,1)-Pt(1,0); S: Pt(1,0)-Pt(0,0)
    sW, fwx, fwy = edge_geo(X0, Y0, X0, Y1, nxS, nyS, phiS)
    sN, fnx, fny = edge_geo(X0, Y1, X1, Y1, nxS, nyS, phiS)
    sE, fex, fey = edge_geo(X1, Y1, X1, Y0, nxS, nyS, phiS)
    sS, fsx, fsy = edge_geo(X1, Y0, X0, Y0, nxS, nyS, phiS)
    ex = (sE - sW) * DX
    ny_ = (sN - sS) * DY
    WlLh = np.sqrt(ex * ex + ny_ * ny_)
    # face-centre ghosts from the raw centres (:91-117)
    fcex, fcey, fcnx, fcny = (_ghosted(nj, ni) for _ in range(4))
    fcex[1:-1, 1:-1] = fex
    fcey[1:-1, 1:-1] = fey
    fcnx[1:-1, 1:-1] = fnx
    fcny[1:-1, 1:-1] = fny
    J = slice(1, nj + 1)
    fcex[J, 0] = fcex[J, 1] - dx[0]
    fcey[J, 0] = fcey[J, 1]
    fcex[J, ni + 1] = fcex[J, ni] + dx[ni - 1]
    fcey[J, ni + 1] = fcey[J, ni]
    fcnx[J, 0] = fcnx[J, 1] - dx[0]
    fcny[J, 0] = fcny[J, 1]
    fcnx[J, ni + 1] = fcnx[J, ni] + dx[ni - 1]
    fcny[J, ni + 1] = fcny[J, ni]
    fcex[0, :] = fcex[1, :]
    fcey[0, :] = fcey[1, :] - dy[0]
    fcex[nj + 1, :] = fcex[nj, :]
    fcey[nj + 1, :] = fcey[nj, :] + dy[nj - 1]
    fcnx[0, :] = fcex[1, :]
    fcny[0, :] = fcey[1, :] - dy[0]
    fcnx[nj + 1, :] = fcex[nj, :]
    fcny[nj + 1, :] = fcey[nj, :] + dy[nj - 1]
    # fluid fractions and fluid face centres (:119-159)
    aW, aE, aN, aS = 1.0 - sW, 1.0 - sE, 1.0 - sN, 1.0 - sS

    def conv(c, a, h):
        out = np.where(np.abs(c) < EPSI, 0.0, (0.5 - a / 2.0) * np.copysign(1.0, -c) * h)
        return np.where(a < EPSIF, 0.0, out)

    fey = conv(fey, aE, DY)
    fwy = conv(fwy, aW, DY)
    fnx = conv(fnx, aN, DX)
    fsx = conv(fsx, aS, DX)
    # harmonic averaging over i = 2..Isize-1, j = 2..Jsize-1 (:161-204):
    # cell slices [1:nj-1, 1:ni-1]
    if ni >= 3 and nj >= 3:
        c = (slice(1, nj - 1), slice(1, ni - 1))
        up = (slice(2, nj), slice(1, ni - 1))
        rt = (slice(1, nj - 1), slice(2, ni))
        n_, s_ = aN[c].copy(), aS[up].copy()
        den = n_ + s_ + ROUNDOFF
        avg = 2.0 * n_ * s_ / den
        yd = (y[2:nj] - y[1:nj - 1])[:, None]
        nfx = (fnx[c] * n_ + fsx[up] * s_) / den
        nfy = (fny[c] * n_ + (fsy[up] + yd) * s_) / den
        nfy = np.where((n_ < EPSIF) & (s_ < EPSIF), DY[c] / 2.0, nfy)
        fnx[c] = nfx
        fny[c] = nfy
        aN[c] = avg
        aS[up] = avg
        e_, w_ = aE[c].copy(), aW[rt].copy()
        den = e_ + w_ + ROUNDOFF
        avg = 2.0 * e_ * w_ / den
        xd = (x[2:ni] - x[1:ni - 1])[None, :]
        efx = (fex[c] * e_ + (fwx[rt] + xd) * w_) / den
        efy = (fey[c] * e_ + fwy[rt] * w_) / den
        efx = np.where((e_ < EPSIF) & (w_ < EPSIF), DX[c] / 2.0, efx)
        fex[c] = efx
        fey[c] = efy
        aE[c] = avg
        aW[rt] = avg
    fcex[1:-1, 1:-1] = fex
    fcey[1:-1, 1:-1] = fey
    fcnx[1:-1, 1:-1] = fnx
    fcny[1:-1, 1:-1] = fny
    # closings (:214-243)
    solid0 = vofS > 1.0 - EPSI
    closed = solid0.copy()
    if p_cell:
        fluid0 = ~solid0
        wsum = aE      # E and N are the averaged values when (i,j) is visited
        while True:
            wl = np.zeros((nj, ni), dtype=bool)
            wl[:, 1:] = closed[:, :-1]
            sl = np.zeros((nj, ni), dtype=bool)
            sl[1:, :] = closed[:-1, :]
            Wv = np.where(wl, 0.0, aW)
            Sv = np.where(sl, 0.0, aS)
            new = solid0 | (fluid0 & ((Wv + wsum < ROUNDOFF) | (Sv + aN < ROUNDOFF)))
            if np.array_equal(new, closed):
                break
            closed = new
    vofS = np.where(closed, 1.0, vofS)
    gE, gW, gN, gS = (_ghosted(nj, ni) for _ in range(4))
    gE[1:-1, 1:-1] = aE
    gW[1:-1, 1:-1] = aW
    gN[1:-1, 1:-1] = aN
    gS[1:-1, 1:-1] = aS
    cj, ci = np.nonzero(closed)
    gj, gi = cj + 1, ci + 1
    for arr in (gE, gW, gN, gS):
        arr[gj, gi] = 0.0
    gE[gj, gi - 1] = 0.0
    gN[gj - 1, gi] = 0.0
    gW[gj, gi + 1] = 0.0
    gS[gj + 1, gi] = 0.0
    # wall length of full solid cells (:245-258)
    full = np.abs(vofS - 1.0) < ROUNDOFF
    ew0 = (gW[1:-1, 1:-1] == 0.0) | (gE[1:-1, 1:-1] == 0.0)
    ns0 = (gN[1:-1, 1:-1] == 0.0) | (gS[1:-1, 1:-1] == 0.0)
    WlLh = np.where(full & ew0, DY, WlLh)
    WlLh = np.where(full & ns0, DX, WlLh)
    nxS = np.array(nxS, dtype=np.float64)
    nyS = np.array(nyS, dtype=np.float64)
    # corner patch (:259-286), P cells, j = 1..Jsize-4
    if p_cell and icorprob and nj > 4:
        s0 = vofS[:nj - 4]
        s1 = vofS[1:nj - 3]
        s2 = vofS[2:nj - 2]
        low = (s0 > -EPSI) & (s0 < EPSI)
        bar = low & (s1 > 0.6 - EPSI) & (s1 < 0.6 + EPSI) & (s2 > 0.6 - EPSI) & (s2 < 0.6 + EPSI)
        flat = low & (s1 > 1.0 - EPSI) & (s1 < 1.0 + EPSI)
        bj, bi = np.nonzero(bar)
        gN[bj + 1, bi + 1] = 0.4  # the reference's 0.40 is default real; 0.4 keeps the face symmetric
        gS[bj + 2, bi + 1] = 0.4
        fcnx[bj + 1, bi + 1] = -0.3 * dx[bi]
        fcny[bj + 1, bi + 1] = 0.5 * dy[bj]
        WlLh[bj, bi] = 0.6 * dx[bi]
        nyS[bj, bi] = -1.0
        nxS[bj, bi] = 0.0
        fj, fi = np.nonzero(flat)
        WlLh[fj, fi] = dx[fi]
        nyS[fj, fi] = -1.0
        nxS[fj, fi] = 0.0
    return {"aE": gE, "aW": gW, "aN": gN, "aS": gS, "fcex": fcex, "fcey": fcey, "fcnx": fcnx, "fcny": fcny,
            "WlLh": WlLh, "vofS": vofS, "nxS": nxS, "nyS": nyS}


def define_moex(P, U, V):
    """DefineMomentumExchangeCell (Cutcell_Mod.f90:360-397); each cell reads
    only the P faces it writes itself, so the cells are independent.  Closes
    P faces in place; returns (moexU, moexV)."""
    vS, vU, vV = P["vofS"], U["vofS"], V["vofS"]
    nj, ni = vS.shape
    aE, aW, aN, aS = P["aE"], P["aW"], P["aN"], P["aS"]
    solid = vS >= 1.0 - EPSI
    cu = (vU >= EPSI) & (vU < 1.0 - EPSI)
    cv = (vV >= EPSI) & (vV < 1.0 - EPSI)
    mu = np.zeros((nj, ni), np.int32)
    mv = np.zeros((nj, ni), np.int32)
    inner_u = cu[:, :-1] & (solid[:, :-1] | solid[:, 1:] | (aE[1:-1, 1:-2] < EPSIF))
    mu[:, :-1] = inner_u
    mu[:, -1] = cu[:, -1] & solid[:, -1]
    inner_v = cv[:-1, :] & (solid[:-1, :] | solid[1:, :] | (aN[1:-2, 1:-1] < EPSIF))
    mv[:-1, :] = inner_v
    mv[-1, :] = cv[-1, :] & solid[-1, :]
    uj, ui = np.nonzero(inner_u)
    aE[uj + 1, ui + 1] = 0.0
    aW[uj + 1, ui + 2] = 0.0
    vj, vi = np.nonzero(inner_v)
    aN[vj + 1, vi + 1] = 0.0
    aS[vj + 2, vi + 1] = 0.0
    return mu, mv


def posnu_active(vofS, iu, iv):
    """NumberExternalCell (Cutcell_Mod.f90:340-358): (numbered, active).
    numbered: Posnu /= -1, i.e. fluid inside i <= Isize-iu, j <= Jsize-iv,
    and every cell of the column / row the loop does not reach (their Posnu
    keeps a fresh allocation's 0); active: a row of the momentum system
    (fluid and reached)."""
    fluid = vofS < 1.0 - EPSI
    nj, ni = vofS.shape
    numbered = fluid.copy()
    act = fluid.copy()
    if iu:
        numbered[:, ni - 1] = True
        act[:, ni - 1] = False
    if iv:
        numbered[nj - 1, :] = True
        act[nj - 1, :] = False
    return numbered.astype(np.int32), act.astype(np.int32)


def _fill_unwritten(a, east):
    if east:
        a[1:-1, -1] = a[1:-1, -2]
        a[0, :] = a[1, :]
        a[-1, :] = a[-2, :]
    else:
        a[-1, 1:-1] = a[-2, 1:-1]
        a[:, 0] = a[:, 1]
        a[:, -1] = a[:, -2]


def _east(T, TG, bdx, prevAlN1):
    """EastFaceInterpolationInf (Cutcell_Mod.f90:551-616); bdx = BGrid%dx
    of the reference's column i+iu, per i = 1..Isize-1 (length ni-1)."""
    ccx, ccy, act = T["ccx"], T["ccy"], T["numbered"]
    nxS, nyS, phiS = T["nxS"], T["nyS"], T["phiS"]
    nj, ni = ccx.shape
    L = (slice(None), slice(0, ni - 1))
    R = (slice(None), slice(1, ni))
    bd = np.broadcast_to(bdx[None, :], (nj, ni - 1))
    pair = (act[L] != 0) & (act[R] != 0)
    fex = T["fcex"][1:-1, 1:ni]
    fey = T["fcey"][1:-1, 1:ni]
    with np.errstate(invalid="ignore", divide="ignore"):
        SxE = ccx[R] + bd - ccx[L]
        Sy = ccy[R] - ccy[L]
        eta = np.abs(fex - ccx[L]) / SxE
        eta = np.where(eta >= 1.0, 0.5, eta)
        xf = (1.0 - eta) * ccx[L] + eta * (ccx[R] + bd)
        yf = (1.0 - eta) * ccy[L] + eta * ccy[R]
        nyf = 0.5 * (nyS[L] + nyS[R])
        ipd = 0.5 * (np.abs(xf * nxS[L] + yf * nyS[L] + phiS[L]) +
                     np.abs((xf - bd) * nxS[R] + yf * nyS[R] + phiS[R]))
        fcd = 0.5 * (np.abs(fex * nxS[L] + fey * nyS[L] + phiS[L]) +
                     np.abs((fex - bd) * nxS[R] + fey * nyS[R] + phiS[R]))
        AlE = np.abs(fcd) / np.abs(ipd)
        DAlE = Sy * nyf / (SxE * np.abs(ipd))
    small = np.abs(ipd) < 1.0e-5
    DAlE = np.where(small, 0.0, DAlE)
    AlE = np.where(small, 1.0, AlE)
    AlE = np.where(AlE >= 2.0, 2.0, AlE)
    tilt = np.abs(Sy) > 0.0001 * np.broadcast_to(TG[3][:, None], Sy.shape)
    DAlE = np.where(tilt, DAlE, 0.0)
    AlE = np.where(tilt, AlE, 1.0)
    out = {}
    out["SxE"] = np.where(pair, SxE, bd)
    out["EtaE"] = np.where(pair, eta, 0.5)
    out["AlE"] = np.where(pair, AlE, 1.0)
    out["DAlE"] = np.where(pair, DAlE, 0.0)
    res = {}
    for k in ("SxE", "EtaE", "AlE", "DAlE"):
        a = _ghosted(nj, ni)
        a[1:-1, 1:ni] = out[k]
        a[1:-1, ni] = a[1:-1, ni - 1]
        a[1:-1, 0] = a[1:-1, 1] if k != "AlE" else prevAlN1
        _fill_unwritten(a, True)
        res[k] = a
    return res


def _north(T, TG, bdy):
    """NorthFaceInterpolationInf (Cutcell_Mod.f90:618-700); bdy = BGrid%dy
    of the reference's row j+iv, per j = 1..Jsize-1 (length nj-1).  Its
    factor is IntPointDist / FaceCenterDist, the east one's inverse."""
    ccx, ccy, act = T["ccx"], T["ccy"], T["numbered"]
    nxS, nyS, phiS = T["nxS"], T["nyS"], T["phiS"]
    nj, ni = ccx.shape
    B = (slice(0, nj - 1), slice(None))
    A = (slice(1, nj), slice(None))
    bd = np.broadcast_to(bdy[:, None], (nj - 1, ni))
    pair = (act[B] != 0) & (act[A] != 0)
    fnx = T["fcnx"][1:nj, 1:-1]
    fny = T["fcny"][1:nj, 1:-1]
    with np.errstate(invalid="ignore", divide="ignore"):
        SyN = ccy[A] + bd - ccy[B]
        Sx = ccx[A] - ccx[B]
        eta = np.abs(fny - ccy[B]) / SyN
        eta = np.where(eta >= 1.0, 0.5, eta)
        xf = (1.0 - eta) * ccx[B] + eta * ccx[A]
        yf = (1.0 - eta) * ccy[B] + eta * (ccy[A] + bd)
        nxf = 0.5 * (nxS[B] + nxS[A])
        ipd = 0.5 * (np.abs(xf * nxS[B] + yf * nyS[B] + phiS[B]) +
                     np.abs(xf * nxS[A] + (yf - bd) * nyS[A] + phiS[A]))
        fcd = 0.5 * (np.abs(fnx * nxS[B] + fny * nyS[B] + phiS[B]) +
                     np.abs(fnx * nxS[A] + (fny - bd) * nyS[A] + phiS[A]))
        AlN = np.abs(ipd) / np.abs(fcd)
        DAlN = Sx * nxf / (SyN * np.abs(ipd))
    small = np.abs(ipd) < 1.0e-5
    DAlN = np.where(small, 0.0, DAlN)
    AlN = np.where(small, 1.0, AlN)
    AlN = np.where(AlN >= 2.0, 2.0, AlN)
    tilt = np.abs(Sx) >= 0.0001 * np.broadcast_to(TG[2][None, :], Sx.shape)
    DAlN = np.where(tilt, DAlN, 0.0)
    AlN = np.where(tilt, AlN, 1.0)
    out = {"SyN": np.where(pair, SyN, bd), "EtaN": np.where(pair, eta, 0.5), "AlN": np.where(pair, AlN, 1.0),
           "DAlN": np.where(pair, DAlN, 0.0)}
    res = {}
    for k in ("SyN", "EtaN", "AlN", "DAlN"):
        a = _ghosted(nj, ni)
        a[1:nj, 1:-1] = out[k]
        a[nj, 1:-1] = a[nj - 1, 1:-1]
        a[0, 1:-1] = a[1, 1:-1]
        _fill_unwritten(a, False)
        res[k] = a
    return res


def new_cell_face(g, axes, P, U, V, prev_AlN=None):
    """NewCellFace (Cutcell_Mod.f90:399-549) and the interpolation factors
    (:545-548) for the U and V cells, in place on the dicts U and V."""
    PDX = np.broadcast_to(g.dx[None, :], P["vofS"].shape)
    PDY = np.broadcast_to(g.dy[:, None], P["vofS"].shape)
    nj, ni = PDX.shape
    for T, k in ((U, 0), (V, 1)):
        T["ccx"] = np.zeros((nj, ni))
        T["ccy"] = np.zeros((nj, ni))
        big = (T["moex"] != 1) & (T["vofS"] > EPSI)
        if k == 0:
            T["ccx"] = np.where(big, P["fcex"][1:-1, 1:-1] - PDX / 2.0, 0.0)
            T["ccy"] = np.where(big, P["fcey"][1:-1, 1:-1], 0.0)
        else:
            T["ccx"] = np.where(big, P["fcnx"][1:-1, 1:-1], 0.0)
            T["ccy"] = np.where(big, P["fcny"][1:-1, 1:-1] - PDY / 2.0, 0.0)
        # small cells next to one P cell (:427-489)
        small = (T["moex"] == 1) & (T["vofS"] < 1.0 - EPSI)
        if np.any(small):
            nN = T["aN"][1:-1, 1:-1]
            nS = T["aN"][0:-2, 1:-1]
            eE = T["aE"][1:-1, 1:-1]
            eW = T["aE"][1:-1, 0:-2]
            fnN = T["fcnx"][1:-1, 1:-1]
            fnS = T["fcnx"][0:-2, 1:-1]
            feE = T["fcey"][1:-1, 1:-1]
            feW = T["fcey"][1:-1, 0:-2]
            nxS, nyS, phiS = T["nxS"], T["nyS"], T["phiS"]
            cx, cy = T["ccx"].copy(), T["ccy"].copy()
            with np.errstate(invalid="ignore", divide="ignore"):
                ny0 = nN * nS == 0.0
                cx1 = fnN * nN / (nN + TOL) + fnS * nS / (nS + TOL)
                cy1 = feE * eE / (eE + TOL) + feW * eW / (eW + TOL)
                cx2 = -(nyS * cy1 + phiS) / (nxS + TOL)
                cx = np.where(small & ny0, cx1, cx)
                cy = np.where(small & ~ny0, cy1, cy)
                cx = np.where(small & ~ny0, cx2, cx)
                ex0 = eE * eW == 0.0
                cy3 = -(nxS * cx + phiS) / (nyS + TOL)
                cy = np.where(small & ex0, cy1, cy)
                cy = np.where(small & ~ex0, cy3, cy)
            T["ccx"], T["ccy"] = cx, cy
        T["delh"] = np.abs(T["ccx"] * T["nxS"] + T["ccy"] * T["nyS"] + T["phiS"]) + TOL
        T["delh"] = np.where(T["moex"] == 1, 1.0e-20, T["delh"])
        _, _, hx, hy = axes["U" if k == 0 else "V"]
        ex = (T["aE"][1:-1, 1:-1] - T["aW"][1:-1, 1:-1]) * hx[None, :]
        ny_ = (T["aN"][1:-1, 1:-1] - T["aS"][1:-1, 1:-1]) * hy[:, None]
        T["WlLh"] = np.sqrt(ex * ex + ny_ * ny_)
    # U: east against PGrid%dx(i+1), north against VGrid%dy(j); V: east
    # against UGrid%dx(i), north against PGrid%dy(j+1)
    nU = _north(U, axes["U"], g.vdy[:nj - 1])
    nV = _north(V, axes["V"], g.dy[1:])
    for T, nr, bdx in ((U, nU, g.dx[1:]), (V, nV, g.udx[:ni - 1])):
        key = "U" if T is U else "V"
        prev1 = (prev_AlN[key] if prev_AlN is not None else nr["AlN"])[1:-1, 1]
        er = _east(T, axes[key], bdx, prev1)
        T.update(er)
        T.update(nr)


def cut_cell_geometry(g, planes, prev_AlN=None, icorprob=True):
    """Grid_Preprocess + NewCellFace for the solid lines `planes` =
    {"P"|"U"|"V": (phiS, nxS, nyS)} (cell arrays, phiS > 0 in the fluid).
    Returns {"P": dict, "U": dict, "V": dict}."""
    axes = grid_axes(g)
    out = {}
    for k in "PUV":
        phiS, nxS, nyS = (np.array(a, dtype=np.float64) for a in planes[k])
        x, y, dx, dy = axes[k]
        nj, ni = phiS.shape
        DX = np.broadcast_to(dx[None, :], (nj, ni))
        DY = np.broadcast_to(dy[:, None], (nj, ni))
        vofS = solid_fraction(phiS, nxS, nyS, DX, DY)
        d = cell_geo_cal(axes[k], phiS, nxS, nyS, vofS, p_cell=(k == "P"), icorprob=icorprob)
        d["phiS"] = phiS
        out[k] = d
    P, U, V = out["P"], out["U"], out["V"]
    U["moex"], V["moex"] = define_moex(P, U, V)
    U["numbered"], U["active"] = posnu_active(U["vofS"], 1, 0)
    V["numbered"], V["active"] = posnu_active(V["vofS"], 0, 1)
    P["numbered"], P["active"] = posnu_active(P["vofS"], 0, 0)
    new_cell_face(g, axes, P, U, V, prev_AlN)
    return out
