/*
 * CPU FP64 oracle of the coupled level-set / volume-of-fluid transport
 * (Clsvof_Mod.f90).  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and
 * the CPU-baseline leg of bench.py call it as the checker; the product path
 * never links it.
 *
 * Every routine restates the reference subroutine it names, in the
 * reference's loop order, with the reference's operand order in every sum
 * and product (the file is compiled with -ffp-contract=off, so each
 * operation rounds as written).  `x**2.d0` of the reference is x*x.
 * Local variables that the reference leaves holding the previous cell's
 * value (stale values) are kept the same way here; their initial value,
 * undefined in the reference, is 0.
 *
 * Layout: cell arrays a[(i-1)+(j-1)*ni]; ghosted arrays a[i+j*(ni+2)];
 * ue (0:ni,1:nj) at i+(j-1)*(ni+1); ve (1:ni,0:nj) at (i-1)+j*ni.
 *
 * Pinned by independent checks (tests/test_oracle_clsvof.py): orc_frac,
 * orc_cellgeocal (polygon clipping), orc_plic_distance and
 * orc_distance_fluid (round trips), orc_reconstruct on straight lines,
 * orc_redistance (straight line, circle), the sweeps (exact translation,
 * uniform states, volume; the level-set face value by exact translation of a
 * linear profile and first-order upwind differences on a quadratic one, both
 * sweeps and both flow directions; the one-sided level-set fluxes of the
 * first and last faces and the inflow at face 0 by the positions they sample
 * on a linear profile, uniform and stretched grids), orc_inflow_level (the
 * sine's known values at the west edge) and the rows the west inflow enters,
 * orc_interface_to_uv (straight lines, solid bounds).  Parity unpinned
 * (restated line by line, nothing independent fixes them): the
 * quality-based stencil choice of orc_normal_irre, the stale normals on the
 * domain edge, the tie order of orc_redistance, and the stale-value
 * walk-back of orc_interface_to_uv.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sg_oracle.h"

#define VOFEPS 1.0e-14  /* Clsvof_Mod.f90:10 */
#define TOLDENO 1.0e-24 /* Clsvof_Mod.f90:10 */
#define EPSI 1.0e-3     /* StateVariables.f90:28 */
#define BAND_WIDTH 4    /* Clsvof_Mod.f90:9 */

#define C(a, i, j) (a)[((i)-1) + (size_t)((j)-1) * (size_t)g->ni]
#define GH(a, i, j) (a)[(i) + (size_t)(j) * (size_t)(g->ni + 2)]

static double sgn(double a, double b) { return copysign(fabs(a), b); } /* Fortran SIGN */

/* frac, Clsvof_Mod.f90:1371-1404: liquid fraction of a dx*dy rectangle
 * centred at the origin below the line |nx| x + |ny| y + d = 0. */
double orc_frac(double nx, double ny, double diss, double dx, double dy)
{
    const double slop_eps = 1.0e-20;
    double tnx = fabs(nx), tny = fabs(ny);
    if (diss + 0.5 * (tnx * dx + tny * dy) <= 0.0) return 1.0;
    if (diss - 0.5 * (tnx * dx + tny * dy) >= 0.0) return 0.0;
    if (tnx <= slop_eps) return 0.5 - diss / dy;
    if (tny <= slop_eps) return 0.5 - diss / dx;
    double xx = (0.5 * tny * dy - diss) / dx / tnx + 0.5;
    double yy = (0.5 * tnx * dx - diss) / dy / tny + 0.5;
    double totalarea = 0.5 * xx * yy;
    double topvf = 0.0;
    if (yy > 1.0) {
        double t = (yy - 1.0) / yy;
        topvf = t * t;
    }
    double rightvf = 0.0;
    if (xx > 1.0) {
        double t = (xx - 1.0) / xx;
        rightvf = t * t;
    }
    return totalarea * (1.0 - topvf - rightvf);
}

/* distance, Clsvof_Mod.f90:1323-1369: plane constant of the PLIC line with
 * normal (nxx,nyy) that cuts volume fraction vf from a dx*dy cell. */
double orc_plic_distance(double vf, double nxx, double nyy, double dx, double dy)
{
    const double slop_eps = 1.0e-20, equal_eps = 1.0e-20;
    double nx = fabs(nxx), ny = fabs(nyy);
    if (nx < slop_eps) return (0.5 - vf) * dy;
    if (ny < slop_eps) return (0.5 - vf) * dx;
    double dll = 0.5 * (nx * dx + ny * dy);
    double dul = 0.5 * (nx * dx - ny * dy);
    double dlr = -dul;
    double dur = -dll;
    double vful = orc_frac(nx, ny, dul, dx, dy);
    double vflr = orc_frac(nx, ny, dlr, dx, dy);
    double t;
    if (vf <= vful && vf <= vflr) {
        if (vful > vflr) {
            t = vful; vful = vflr; vflr = t;
            t = dul; dul = dlr; dlr = t;
        }
        double d = dul - dll;
        return dll - sqrt((d * d) * vf / vful);
    } else if (vf >= vful && vf >= vflr) {
        if (vflr < vful) {
            t = vful; vful = vflr; vflr = t;
            t = dul; dul = dlr; dlr = t;
        }
        double d = dlr - dur;
        return dur + sqrt(d * d * (1.0 - vf) / (1.0 - vflr));
    } else {
        if (fabs(vflr - vful) < equal_eps) return 0.5 * (dul + dlr);
        double theta = (vf - vful) / (vflr - vful);
        return dlr * theta + dul * (1.0 - theta);
    }
}

/* Fluid polygon of a cell cut by the solid line (nodes with
 * nxs x + nys y + phis >= 0, walked counter-clockwise from the SE corner),
 * shared construction of CellGeoCal :2170-2219 and DistanceFluidCalculate
 * :1966-2015.  node holds 6 points; returns temp (the index, 1-based, of the
 * closing copy of node 1). */
static int solid_polygon(double nxs, double nys, double phis, double dx, double dy, double node[6][2])
{
    double dx2 = dx / 2.0, dy2 = dy / 2.0;
    int temp = 1;
    memset(node, 0, sizeof(double) * 12);
    double dpt[4];
    dpt[0] = dx2 * nxs - dy2 * nys + phis;
    dpt[1] = -dx2 * nxs - dy2 * nys + phis;
    dpt[2] = -dx2 * nxs + dy2 * nys + phis;
    dpt[3] = dx2 * nxs + dy2 * nys + phis;
    /* at most 5 vertices: a convex quadrilateral cut by one line (the
     * bound keeps a rounding-degenerate sign pattern inside node) */
#define PUT(a, b) do { if (temp <= 5) { node[temp - 1][0] = (a); node[temp - 1][1] = (b); temp++; } } while (0)
    if (dpt[0] >= 0.0) PUT(dx2, -dy2);
    if (dpt[0] * dpt[1] < 0.0) PUT((dy2 * nys - phis) / nxs, -dy2);
    if (dpt[1] >= 0.0) PUT(-dx2, -dy2);
    if (dpt[1] * dpt[2] < 0.0) PUT(-dx2, (dx2 * nxs - phis) / nys);
    if (dpt[2] >= 0.0) PUT(-dx2, dy2);
    if (dpt[2] * dpt[3] < 0.0) PUT((-dy2 * nys - phis) / nxs, dy2);
    if (dpt[3] >= 0.0) PUT(dx2, dy2);
    if (dpt[3] * dpt[0] < 0.0) PUT(dx2, (-dx2 * nxs - phis) / nys);
#undef PUT
    node[temp - 1][0] = node[0][0];
    node[temp - 1][1] = node[0][1];
    return temp;
}

/* Liquid part (pdt <= 0) of the closed polygon node[0..temp-1] and its
 * signed shoelace area, CellGeoCal :2227-2251 / LiqFuncVol :2133-2153. */
static double liquid_area(const double node[6][2], int temp, const double* pdt)
{
    double nodel[7][2];
    int templ = 1;
    for (int k = 1; k <= temp - 1; k++) {
        if (pdt[k - 1] <= 0.0) {
            nodel[templ - 1][0] = node[k - 1][0];
            nodel[templ - 1][1] = node[k - 1][1];
            templ++;
        }
        if (pdt[k - 1] * pdt[k] < 0.0) {
            double a = fabs(pdt[k - 1]), b = fabs(pdt[k]);
            nodel[templ - 1][0] = node[k - 1][0] + (node[k][0] - node[k - 1][0]) * a / (a + b);
            nodel[templ - 1][1] = node[k - 1][1] + (node[k][1] - node[k - 1][1]) * a / (a + b);
            templ++;
        }
    }
    if (templ > 1) {
        nodel[templ - 1][0] = nodel[0][0];
        nodel[templ - 1][1] = nodel[0][1];
    }
    double vol = 0.0;
    for (int k = 1; k <= templ - 1; k++)
        vol = vol + 0.5 * (nodel[k - 1][0] * nodel[k][1] - nodel[k][0] * nodel[k - 1][1]);
    return vol;
}

/* CellGeoCal, Clsvof_Mod.f90:2159-2253: solid fraction vols and liquid
 * fraction volf of a dx*dy cell cut by the solid line (nxs,nys,phis) and
 * the liquid line (nxl,nyl,phil). */
void orc_cellgeocal(double nxs, double nys, double phis, double nxl, double nyl, double phil,
                    double dx, double dy, double* vols, double* volf)
{
    double node[6][2], pdt[6] = {0.0};
    int temp = solid_polygon(nxs, nys, phis, dx, dy, node);
    double vol = 0.0;
    for (int k = 1; k <= temp - 1; k++) {
        vol = vol + 0.5 * (node[k - 1][0] * node[k][1] - node[k][0] * node[k - 1][1]);
        pdt[k - 1] = node[k - 1][0] * nxl + node[k - 1][1] * nyl + phil;
    }
    *vols = 1.0 - fabs(vol / (dx * dy));
    pdt[temp - 1] = pdt[0];
    vol = liquid_area(node, temp, pdt);
    *volf = fabs(vol / (dx * dy));
}

/* LiqFuncVol, Clsvof_Mod.f90:2107-2155: volf minus the liquid fraction left
 * in the fluid polygon by the line through xf with normal (nxf,nyf). */
static double liq_func_vol(double nxf, double nyf, double volf, double dx, double dy,
                           const double node[6][2], const double xf[2])
{
    double phil = -(nxf * xf[0] + nyf * xf[1]);
    int temp = 0;
    for (int k = 1; k <= 6; k++)
        if (fabs(node[k - 1][0]) > 1.0e-14 || fabs(node[k - 1][1]) > 1.0e-14) temp = k;
    double pdt[7] = {0.0};
    for (int k = 1; k <= temp - 1; k++) pdt[k - 1] = node[k - 1][0] * nxf + node[k - 1][1] * nyf + phil;
    if (temp == 0) return volf; /* reference pauses here (:2124-2128) */
    pdt[temp - 1] = pdt[0];
    double vol = liquid_area(node, temp, pdt);
    return volf - fabs(vol / (dx * dy));
}

/* zriddr, Clsvof_Mod.f90:2044-2105: Ridders' root search along the segment
 * x1 -> x2 for the line position that leaves volf of liquid.  Returns 1 and
 * sets *phif when it assigns phif, 0 when the reference leaves it alone. */
static int zriddr(double nxf, double nyf, double volf, double dx, double dy, double* phif,
                  const double node[6][2], const double x1[2], const double x2[2], double tol, int itmax)
{
    const double UNUSED = -1.11e-30;
    double fl = liq_func_vol(nxf, nyf, volf, dx, dy, node, x1);
    double fh = liq_func_vol(nxf, nyf, volf, dx, dy, node, x2);
    if ((fl > tol && fh < -tol) || (fl < -tol && fh > tol)) {
        double xl[2] = {x1[0], x1[1]}, xh[2] = {x2[0], x2[1]};
        double ans[2] = {UNUSED, UNUSED}, xm[2], xnew[2];
        for (int iter = 1; iter <= itmax; iter++) {
            xm[0] = 0.5 * (xl[0] + xh[0]);
            xm[1] = 0.5 * (xl[1] + xh[1]);
            double fm = liq_func_vol(nxf, nyf, volf, dx, dy, node, xm);
            double s = sqrt(fm * fm - fl * fh);
            double sg = sgn(1.0, fl - fh);
            xnew[0] = xm[0] + (xm[0] - xl[0]) * sg * fm / s;
            xnew[1] = xm[1] + (xm[1] - xl[1]) * sg * fm / s;
            double e0 = xnew[0] - ans[0], e1 = xnew[1] - ans[1];
            if (sqrt(e0 * e0 + e1 * e1) <= tol) {
                *phif = -(nxf * ans[0] + nyf * ans[1]);
                return 1;
            }
            ans[0] = xnew[0];
            ans[1] = xnew[1];
            double fnew = liq_func_vol(nxf, nyf, volf, dx, dy, node, ans);
            if (fabs(fnew) <= tol) {
                *phif = -(nxf * ans[0] + nyf * ans[1]);
                return 1;
            }
            if (sgn(fm, fnew) != fm) {
                xl[0] = xm[0]; xl[1] = xm[1];
                fl = fm;
                xh[0] = ans[0]; xh[1] = ans[1];
                fh = fnew;
            } else if (sgn(fl, fnew) != fl) {
                xh[0] = ans[0]; xh[1] = ans[1];
                fh = fnew;
            } else if (sgn(fh, fnew) != fh) {
                xl[0] = ans[0]; xl[1] = ans[1];
                fl = fnew;
            }
            double h0 = xh[0] - xl[0], h1 = xh[1] - xl[1];
            if (sqrt(h0 * h0 + h1 * h1) <= tol) {
                *phif = -(nxf * ans[0] + nyf * ans[1]);
                return 1;
            }
        }
        *phif = -(nxf * ans[0] + nyf * ans[1]);
        return 1;
    }
    if (fl < tol && fl > -tol) {
        *phif = -(nxf * x1[0] + nyf * x1[1]);
        return 1;
    }
    if (fh < tol && fh > -tol) {
        *phif = -(nxf * x2[0] + nyf * x2[1]);
        return 1;
    }
    return 0;
}

/* DistanceFluidCalculate, Clsvof_Mod.f90:1957-2042: plane constant of the
 * liquid line with normal (nxf,nyf) in a cell cut by the solid line.
 * Returns 0 (and leaves *phif) when the reference leaves it unassigned. */
int orc_distance_fluid(double nxss, double nyss, double phiss, double nxf, double nyf, double volf,
                       double dx, double dy, double* phif)
{
    double node[6][2];
    int temp = solid_polygon(nxss, nyss, phiss, dx, dy, node);
    double diag = sqrt(dx * dx + dy * dy);
    double maxdis = -2.0 * diag, mindis = 2.0 * diag;
    double maxver[2] = {0.0, 0.0}, minver[2] = {0.0, 0.0};
    for (int k = 1; k <= temp - 1; k++) {
        double dt23 = nxf * node[k - 1][0] + nyf * node[k - 1][1];
        if (dt23 > maxdis) {
            maxdis = dt23;
            maxver[0] = node[k - 1][0];
            maxver[1] = node[k - 1][1];
        }
        if (dt23 < mindis) {
            mindis = dt23;
            minver[0] = node[k - 1][0];
            minver[1] = node[k - 1][1];
        }
    }
    maxver[0] = maxver[0] + maxver[0] * nxf * EPSI;
    maxver[1] = maxver[1] + maxver[1] * nyf * EPSI;
    minver[0] = minver[0] - minver[0] * nxf * EPSI;
    minver[1] = minver[1] - minver[1] * nyf * EPSI;
    return zriddr(nxf, nyf, volf, dx, dy, phif, node, minver, maxver, 1.0e-14, 1000);
}

/* ---- normals ------------------------------------------------------------ */

/* Vset as Coupled_LS_VOF leaves it (:746-753): the second statement of
 * line 747 overwrites Vset(1,2) with +1 and Vset(2,2) is never assigned
 * (zero module storage), so direction 2 is (0,0) and never selected. */
static const double VSET[8][2] = {{0.0, 1.0}, {0.0, 0.0}, {-1.0, 0.0}, {-1.0, -1.0},
                                  {-1.0, 1.0}, {1.0, 0.0}, {1.0, -1.0}, {1.0, 1.0}};

typedef struct {
    const orc_cls_geom* g;
    const double* phi;
    const double* vofS;
} nv_ctx;

#define PHI(i, j) C(c->phi, i, j)
#define XC(i) (g->xc[(i)-1])
#define YC(j) (g->yc[(j)-1])

/* the "most parallel direction" one-sided pair of Normal_Vector_Irre
 * (:1768-1827 / :1855-1905) */
static void nv_directional(const nv_ctx* c, int i, int j, int imv_init, double* nx, double* ny)
{
    const orc_cls_geom* g = c->g;
    const double eta = 0.075;
    double nuy = eta / 8.0;
    double vx = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
    double vy = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
    double maxvect = 0.0;
    int imv = imv_init;
    for (int ii = 1; ii <= 8; ii++) {
        double a = VSET[ii - 1][0], b = VSET[ii - 1][1];
        double cospar = fabs(a * vx + b * vy) / (sqrt(a * a + b * b) * sqrt(vx * vx + vy * vy));
        if (cospar >= maxvect) {
            maxvect = cospar;
            imv = ii;
        }
    }
    double v1x, v1y;
    if (imv == 0) {
        /* the reference reads outside Vset here (zero gradient); taken as (-1,0) */
        v1x = -1.0;
        v1y = 0.0;
    } else {
        v1x = -VSET[imv - 1][1];
        v1y = VSET[imv - 1][0];
    }
    double vx1, vx2, vy1, vy2;
    if (v1x == -1.0) {
        vx1 = (PHI(i, j) - PHI(i - 1, j)) / (XC(i) - XC(i - 1));
        vx2 = (PHI(i + 1, j) - PHI(i, j)) / (XC(i + 1) - XC(i));
    } else if (v1x == 0.0) {
        vx1 = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
        vx2 = vx1;
    } else {
        vx1 = (PHI(i + 1, j) - PHI(i, j)) / (XC(i + 1) - XC(i));
        vx2 = (PHI(i, j) - PHI(i - 1, j)) / (XC(i) - XC(i - 1));
    }
    if (v1y == -1.0) {
        vy1 = (PHI(i, j) - PHI(i, j - 1)) / (YC(j) - YC(j - 1));
        vy2 = (PHI(i, j + 1) - PHI(i, j)) / (YC(j + 1) - YC(j));
    } else if (v1y == 0.0) {
        vy1 = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
        vy2 = vy1;
    } else {
        vy1 = (PHI(i, j + 1) - PHI(i, j)) / (YC(j + 1) - YC(j));
        vy2 = (PHI(i, j) - PHI(i, j - 1)) / (YC(j) - YC(j - 1));
    }
    double qv1 = fabs(1.0 - sqrt(vx1 * vx1 + vy1 * vy1));
    double qv2 = fabs(1.0 - sqrt(vx2 * vx2 + vy2 * vy2));
    if (qv1 < qv2 + nuy) {
        *nx = vx1;
        *ny = vy1;
    } else {
        *nx = vx2;
        *ny = vy2;
    }
}

/* Normal_Vector_Irre, Clsvof_Mod.f90:1710-1919 (interior cells only). */
void orc_normal_irre(const orc_cls_geom* g, const double* phi, const double* vofS, int i, int j,
                     double* nx, double* ny)
{
    nv_ctx cc = {g, phi, vofS}, *c = &cc;
    const double eta = 0.075;
    int ni = g->ni, nj = g->nj;
#define VS(a, b) C(vofS, a, b)
    if (VS(i, j) >= EPSI) {
        if (VS(i + 1, j) < 1.0 - EPSI && VS(i - 1, j) < 1.0 - EPSI)
            *nx = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
        else if (VS(i + 1, j) >= 1.0 - EPSI)
            *nx = (PHI(i, j) - PHI(i - 1, j)) / (XC(i) - XC(i - 1));
        else
            *nx = (PHI(i + 1, j) - PHI(i, j)) / (XC(i + 1) - XC(i));
        if (VS(i, j + 1) < 1.0 - EPSI && VS(i, j - 1) < 1.0 - EPSI)
            *ny = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
        else if (VS(i, j + 1) > 1.0 - EPSI)
            *ny = (PHI(i, j) - PHI(i, j - 1)) / (YC(j) - YC(j - 1));
        else
            *ny = (PHI(i, j + 1) - PHI(i, j)) / (YC(j + 1) - YC(j));
        return;
    }
#undef VS
    double vx, vy, qi[3], qj[3];
    if (i >= 3)
        vx = (PHI(i, j) - PHI(i - 2, j)) / (XC(i) - XC(i - 2));
    else
        vx = (PHI(i, j) - PHI(i - 1, j)) / (XC(i) - XC(i - 1));
    vy = (PHI(i - 1, j + 1) - PHI(i - 1, j - 1)) / (YC(j + 1) - YC(j - 1));
    qi[0] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    vx = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
    vy = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
    qi[1] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    if (i <= ni - 2)
        vx = (PHI(i + 2, j) - PHI(i, j)) / (XC(i + 2) - XC(i));
    else
        vx = (PHI(i + 1, j) - PHI(i, j)) / (XC(i + 1) - XC(i));
    vy = (PHI(i + 1, j + 1) - PHI(i + 1, j - 1)) / (YC(j + 1) - YC(j - 1));
    qi[2] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    int ddx;
    if (qi[0] < eta && qi[2] >= eta)
        ddx = -1;
    else if (qi[0] >= eta && qi[2] < eta)
        ddx = 1;
    else if (qi[0] < eta && qi[1] < eta && qi[2] < eta)
        ddx = 0;
    else if (qi[0] >= eta && qi[1] >= eta && qi[2] >= eta)
        ddx = 0;
    else {
        nv_directional(c, i, j, 0, nx, ny);
        return;
    }
    if (j >= 3)
        vy = (PHI(i, j) - PHI(i, j - 2)) / (YC(j) - YC(j - 2));
    else
        vy = (PHI(i, j) - PHI(i, j - 1)) / (YC(j) - YC(j - 1));
    vx = (PHI(i + 1, j - 1) - PHI(i - 1, j - 1)) / (XC(i + 1) - XC(i - 1));
    qj[0] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    vx = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
    vy = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
    qj[1] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    vx = (PHI(i + 1, j + 1) - PHI(i - 1, j + 1)) / (XC(i + 1) - XC(i - 1));
    if (j <= nj - 2)
        vy = (PHI(i, j + 2) - PHI(i, j)) / (YC(j + 2) - YC(j));
    else
        vy = (PHI(i, j + 1) - PHI(i, j)) / (YC(j + 1) - YC(j));
    qj[2] = fabs(1.0 - sqrt(vx * vx + vy * vy));
    int ddy;
    if (qj[0] < eta && qj[2] >= eta)
        ddy = -1;
    else if (qj[0] >= eta && qj[2] < eta)
        ddy = 1;
    else if (qj[0] < eta && qj[1] < eta && qj[2] < eta)
        ddy = 0;
    else if (qj[0] >= eta && qj[1] >= eta && qj[2] >= eta)
        ddy = 0;
    else {
        nv_directional(c, i, j, 1, nx, ny);
        return;
    }
    if (ddx == -1) *nx = (PHI(i, j) - PHI(i - 1, j)) / (XC(i) - XC(i - 1));
    if (ddx == 1) *nx = (PHI(i + 1, j) - PHI(i, j)) / (XC(i + 1) - XC(i));
    if (ddx == 0) *nx = (PHI(i + 1, j) - PHI(i - 1, j)) / (XC(i + 1) - XC(i - 1));
    if (ddy == -1) *ny = (PHI(i, j) - PHI(i, j - 1)) / (YC(j) - YC(j - 1));
    if (ddy == 1) *ny = (PHI(i, j + 1) - PHI(i, j)) / (YC(j + 1) - YC(j));
    if (ddy == 0) *ny = (PHI(i, j + 1) - PHI(i, j - 1)) / (YC(j + 1) - YC(j - 1));
}
#undef PHI

/* Interface_Reconstruct, Clsvof_Mod.f90:1250-1321. */
void orc_reconstruct(const orc_cls_geom* g, const double* vof, const double* phi, double* nx, double* ny,
                     double* dis)
{
    int ni = g->ni, nj = g->nj;
    double nxx = 0.0, nyy = 0.0, diss = 0.0; /* stale across cells, as in the reference */
    for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) {
            double vf = C(vof, i, j), vs = C(g->vofS, i, j);
            int flag = 0;
            if (vf >= VOFEPS && vf < (1.0 - VOFEPS - vs) && vs < EPSI) flag = 1;
            if (vf >= VOFEPS && vf < (1.0 - VOFEPS - vs) && vs >= EPSI && vs < 1.0 - EPSI) flag = 2;
            if (flag) {
                if (i > 1 && i < ni && j > 1 && j < nj) {
                    orc_normal_irre(g, phi, g->vofS, i, j, &nxx, &nyy);
                } else {
                    if (i == 1)
                        nxx = (C(phi, i + 1, j) - C(phi, i, j)) / (XC(i + 1) - XC(i));
                    else if (i == ni)
                        nxx = (C(phi, i, j) - C(phi, i - 1, j)) / (XC(i) - XC(i - 1));
                    if (j == 1)
                        nyy = (C(phi, i, j + 1) - C(phi, i, j)) / (YC(j + 1) - YC(j));
                    else if (j == nj)
                        nyy = (C(phi, i, j) - C(phi, i, j - 1)) / (YC(j) - YC(j - 1));
                }
                double temp = sqrt(nxx * nxx + nyy * nyy);
                if (temp < 1.0e-14) {
                    nxx = 0.0;
                    nyy = 1.0;
                } else {
                    nxx = nxx / temp;
                    nyy = nyy / temp;
                }
                if (flag == 1)
                    diss = orc_plic_distance(vf, nxx, nyy, g->dx[i - 1], g->dy[j - 1]);
                else
                    orc_distance_fluid(C(g->nxS, i, j), C(g->nyS, i, j), C(g->phiS, i, j), nxx, nyy, vf,
                                       g->dx[i - 1], g->dy[j - 1], &diss);
            } else {
                nxx = 0.0;
                nyy = 1.0;
                if (1.0 - vs > VOFEPS)
                    diss = (0.5 - vf / (1.0 - vs)) * g->dy[j - 1];
                else
                    diss = (0.0 - vf) * g->dy[j - 1];
            }
            C(nx, i, j) = nxx;
            C(ny, i, j) = nyy;
            C(dis, i, j) = diss;
        }
    }
}

/* ---- sweeps ---------------------------------------------------------------- */

/* rightflux / leftflux / topflux / bottomflux, Clsvof_Mod.f90:1416-1522.
 * side = +1 right/top (donor cut at the far edge), -1 left/bottom.
 * Width w is the swept length, (a,b) the cell extent along/across the
 * sweep; `along_x` says the sweep is along x. */
static double face_flux(int along_x, int side, double nxx, double nyy, double diss, double volf, double nxss,
                        double nyss, double phiss, double vols, double dx, double dy, double w)
{
    const double eps = 1.0e-14;
    if (w == 0.0) return 0.0;
    double flux;
    double nl = along_x ? nxx : nyy;
    double ns = along_x ? nxss : nyss;
    double len = along_x ? dx : dy;
    double shift = 0.5 * (w - len);
    double dl = side > 0 ? diss - shift * nl : diss + shift * nl;
    if (vols < EPSI) {
        flux = along_x ? orc_frac(nxx, nyy, dl, w, dy) : orc_frac(nxx, nyy, dl, dx, w);
    } else if (vols >= 1.0 - EPSI) {
        return volf;
    } else {
        double ps = side > 0 ? phiss - shift * ns : phiss + shift * ns;
        double volss;
        if (along_x)
            orc_cellgeocal(nxss, nyss, ps, nxx, nyy, dl, w, dy, &volss, &flux);
        else
            orc_cellgeocal(nxss, nyss, ps, nxx, nyy, dl, dx, w, &volss, &flux);
        if (volf / (1.0 - vols) >= 1.0 - VOFEPS) flux = 1.0 - volss;
        flux = flux / (1.0 - volss + TOLDENO);
        if (volf + vols > 1.0 - eps) flux = 1.0;
    }
    return flux * w;
}

/* masked face velocities of Coupled_LS_VOF :778-797 */
void orc_masked_velocities(const orc_cls_geom* g, const double* u, const double* v, double* ue, double* ve)
{
    int ni = g->ni, nj = g->nj;
    for (int j = 1; j <= nj; j++) {
        for (int i = 1; i <= ni; i++)
            ue[i + (size_t)(j - 1) * (ni + 1)] = C(g->uflag, i, j) ? GH(u, i, j) : 0.0;
        ue[0 + (size_t)(j - 1) * (ni + 1)] = GH(u, 0, j);
    }
    for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) ve[(i - 1) + (size_t)j * ni] = C(g->vflag, i, j) ? GH(v, i, j) : 0.0;
        ve[(i - 1) + 0] = GH(v, i, 0);
    }
}

#define UE(i, j) ue[(i) + (size_t)((j)-1) * (ni + 1)]
#define VE(i, j) ve[((i)-1) + (size_t)(j) * ni]

/* X_Sweep, Clsvof_Mod.f90:987-1124 (after its Interface_Reconstruct call,
 * whose nx,ny,dis are passed in).  temvf/temls in/out; `inflow_level` is
 * Depthw+etau of :1057-1059. */
void orc_x_sweep(const orc_cls_geom* g, const double* vfl, const double* phi, const double* nx, const double* ny,
                 const double* dis, const double* ue, double dtv, double inflow_level, double* temvf,
                 double* temls)
{
    int ni = g->ni, nj = g->nj;
    const double* vfls = g->vofS;
    for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++)
            if (C(vfls, i, j) < 1.0 - VOFEPS) C(temvf, i, j) = C(temvf, i, j) / (1.0 - C(vfls, i, j));
    for (int j = 1; j <= nj; j++) {
        double dy = g->dy[j - 1];
        for (int i = 1; i <= ni; i++) {
            double u = UE(i, j), flux;
            if (u >= 0.0) {
                double f = C(vfl, i, j);
                if (f >= (1.0 - VOFEPS) || f <= VOFEPS) {
                    flux = f * u * dtv;
                    if (f >= 1.0) flux = u * dtv;
                    if (f < 0.0) flux = 0.0;
                } else {
                    flux = face_flux(1, 1, C(nx, i, j), C(ny, i, j), C(dis, i, j), f, C(g->nxS, i, j),
                                     C(g->nyS, i, j), C(g->phiS, i, j), C(vfls, i, j), g->dx[i - 1], dy, u * dtv);
                }
                if (GH(g->aE, i, j) > 1.0 - VOFEPS && C(vfls, i, j) > 1.0 - VOFEPS) flux = 1.0 * u * dtv;
            } else {
                if (i < ni) {
                    double f = C(vfl, i + 1, j);
                    if (f >= (1.0 - VOFEPS) || f <= VOFEPS) {
                        flux = f * u * dtv;
                        if (f >= 1.0) flux = u * dtv;
                        if (f < 0.0) flux = 0.0;
                    } else {
                        flux = face_flux(1, -1, C(nx, i + 1, j), C(ny, i + 1, j), C(dis, i + 1, j), f,
                                         C(g->nxS, i + 1, j), C(g->nyS, i + 1, j), C(g->phiS, i + 1, j),
                                         C(vfls, i + 1, j), g->dx[i], dy, -u * dtv);
                        flux = -flux;
                    }
                    if (GH(g->aW, i + 1, j) > 1.0 - VOFEPS && C(vfls, i + 1, j) > 1.0 - VOFEPS)
                        flux = 1.0 * u * dtv;
                } else {
                    flux = C(vfl, i, j) * u * dtv;
                }
            }
            C(temvf, i, j) = C(temvf, i, j) - flux * GH(g->aE, i, j) / g->dx[i - 1];
            if (i <= ni - 1)
                C(temvf, i + 1, j) = C(temvf, i + 1, j) * (1.0 - C(g->vofS, i + 1, j)) +
                                     flux * GH(g->aW, i + 1, j) / g->dx[i];
        }
        double flux;
        if (YC(j) - dy / 2.0 < inflow_level)
            flux = C(vfl, 1, j) * UE(0, j) * dtv;
        else
            flux = 0.0;
        C(temvf, 1, j) = C(temvf, 1, j) + flux * GH(g->aW, 1, j) / g->dx[0];
    }
    /* level set */
    for (int j = 1; j <= nj; j++) {
        for (int i = 2; i <= ni - 1; i++) {
            double u = UE(i, j), lsr;
            if (u >= 0.0)
                lsr = (C(phi, i + 1, j) + C(phi, i - 1, j)) / 2.0;
            else if (i <= ni - 2)
                lsr = (C(phi, i + 2, j) + C(phi, i, j)) / 2.0;
            else
                lsr = (C(phi, i + 1, j) + C(phi, i - 1, j)) / 2.0;
            double flux = u * lsr * dtv;
            C(temls, i, j) = C(temls, i, j) - flux * GH(g->aE, i, j) / g->dx[i - 1];
            C(temls, i + 1, j) = C(temls, i + 1, j) * (1.0 - C(g->vofS, i + 1, j)) +
                                 flux * GH(g->aW, i + 1, j) / g->dx[i];
        }
    }
    for (int j = 1; j <= nj; j++) {
        const double* dx = g->dx;
        double lsr, flux;
        double u1 = UE(1, j);
        if (u1 >= 0.0)
            lsr = C(phi, 1, j) + dx[0] / 2.0 * (1.0 - u1 * dtv / dx[0]) * (C(phi, 2, j) - C(phi, 1, j)) /
                                     (XC(2) - XC(1));
        else
            lsr = C(phi, 2, j) - dx[1] / 2.0 * (1.0 + u1 * dtv / dx[1]) * (C(phi, 3, j) - C(phi, 1, j)) /
                                     (2.0 * dx[1]);
        flux = lsr * u1 * dtv;
        C(temls, 1, j) = C(temls, 1, j) - flux * GH(g->aE, 1, j) / dx[0];
        C(temls, 2, j) = C(temls, 2, j) + flux * GH(g->aE, 1, j) / dx[1];
        C(temls, 1, j) = C(temls, 1, j) + C(phi, 1, j) * UE(0, j) * dtv * GH(g->aW, 1, j) / dx[0];
        double un = UE(ni, j);
        if (un > 0.0)
            lsr = C(phi, ni, j) + dx[ni - 1] / 2.0 * (1.0 - un * dtv / dx[ni - 1]) *
                                      (C(phi, ni, j) - C(phi, ni - 1, j)) / (XC(ni) - XC(ni - 1));
        else
            lsr = C(phi, ni, j) - dx[ni - 2] / 2.0 * (1.0 + un * dtv / dx[ni - 2]) *
                                      (C(phi, ni, j) - C(phi, ni - 1, j)) / (XC(ni) - XC(ni - 1));
        flux = lsr * un * dtv;
        C(temls, ni, j) = C(temls, ni, j) - flux * GH(g->aE, ni, j) / dx[ni - 1];
    }
}

/* Y_Sweep, Clsvof_Mod.f90:1126-1245 (after its Interface_Reconstruct). */
void orc_y_sweep(const orc_cls_geom* g, const double* vfl, const double* phi, const double* nx, const double* ny,
                 const double* dis, const double* ve, double dtv, double* temvf, double* temls)
{
    int ni = g->ni, nj = g->nj;
    const double* vfls = g->vofS;
    for (int j = 1; j <= nj; j++)
        for (int i = 1; i <= ni; i++)
            if (C(vfls, i, j) < 1.0 - VOFEPS) C(temvf, i, j) = C(temvf, i, j) / (1.0 - C(vfls, i, j));
    for (int i = 1; i <= ni; i++) {
        double dx = g->dx[i - 1];
        for (int j = 1; j <= nj; j++) {
            double v = VE(i, j), flux;
            if (v > 0.0) {
                double f = C(vfl, i, j);
                if (f >= (1.0 - VOFEPS) || f <= VOFEPS) {
                    flux = f * v * dtv;
                    if (f >= 1.0) flux = v * dtv;
                    if (f < 0.0) flux = 0.0;
                } else {
                    flux = face_flux(0, 1, C(nx, i, j), C(ny, i, j), C(dis, i, j), f, C(g->nxS, i, j),
                                     C(g->nyS, i, j), C(g->phiS, i, j), C(vfls, i, j), dx, g->dy[j - 1], v * dtv);
                }
                if (GH(g->aN, i, j) > 1.0 - VOFEPS && C(vfls, i, j) > 1.0 - VOFEPS) flux = 1.0 * v * dtv;
            } else {
                if (j < nj) {
                    double f = C(vfl, i, j + 1);
                    if (f >= (1.0 - VOFEPS) || f <= VOFEPS) {
                        flux = f * v * dtv;
                        if (f >= 1.0) flux = v * dtv;
                        if (f < 0.0) flux = 0.0;
                    } else {
                        flux = face_flux(0, -1, C(nx, i, j + 1), C(ny, i, j + 1), C(dis, i, j + 1), f,
                                         C(g->nxS, i, j + 1), C(g->nyS, i, j + 1), C(g->phiS, i, j + 1),
                                         C(vfls, i, j + 1), dx, g->dy[j], -v * dtv);
                        flux = -flux;
                    }
                    if (GH(g->aS, i, j + 1) > 1.0 - VOFEPS && C(vfls, i, j + 1) > 1.0 - VOFEPS)
                        flux = 1.0 * v * dtv;
                } else {
                    flux = C(vfl, i, j) * v * dtv;
                }
            }
            C(temvf, i, j) = C(temvf, i, j) - flux * GH(g->aN, i, j) / g->dy[j - 1];
            if (j <= nj - 1)
                C(temvf, i, j + 1) = C(temvf, i, j + 1) * (1.0 - C(g->vofS, i, j + 1)) +
                                     flux * GH(g->aS, i, j + 1) / g->dy[j];
        }
        C(temvf, i, 1) = C(temvf, i, 1) + C(vfl, i, 1) * VE(i, 0) * dtv * GH(g->aS, i, 1) / g->dy[0];
    }
    for (int j = 2; j <= nj - 1; j++) {
        for (int i = 1; i <= ni; i++) {
            double v = VE(i, j), lst;
            if (v >= 0.0)
                lst = (C(phi, i, j + 1) + C(phi, i, j - 1)) / 2.0;
            else if (j <= nj - 2)
                lst = (C(phi, i, j + 2) + C(phi, i, j)) / 2.0;
            else
                lst = (C(phi, i, j + 1) + C(phi, i, j - 1)) / 2.0;
            double flux = lst * v * dtv;
            C(temls, i, j) = C(temls, i, j) - flux * GH(g->aN, i, j) / g->dy[j - 1];
            C(temls, i, j + 1) = C(temls, i, j + 1) * (1.0 - C(g->vofS, i, j + 1)) +
                                 flux * GH(g->aS, i, j + 1) / g->dy[j];
        }
    }
    for (int i = 1; i <= ni; i++) {
        const double* dy = g->dy;
        double lst, flux;
        double v1 = VE(i, 1);
        if (v1 >= 0.0)
            lst = C(phi, i, 1) + dy[0] / 2.0 * (1.0 - v1 * dtv / dy[0]) * (C(phi, i, 2) - C(phi, i, 1)) /
                                     (YC(2) - YC(1));
        else
            lst = C(phi, i, 2) - dy[1] / 2.0 * (1.0 + v1 * dtv / dy[1]) * (C(phi, i, 3) - C(phi, i, 1)) /
                                     (YC(3) - YC(1));
        flux = lst * v1 * dtv;
        C(temls, i, 1) = C(temls, i, 1) - flux * GH(g->aN, i, 1) / dy[0];
        C(temls, i, 2) = C(temls, i, 2) + flux * GH(g->aN, i, 1) / dy[1];
        C(temls, i, 1) = C(temls, i, 1) + C(phi, i, 1) * VE(i, 0) * dtv * GH(g->aS, i, 1) / dy[0];
        double vn = VE(i, nj);
        if (vn > 0.0)
            lst = C(phi, i, nj) + dy[nj - 1] / 2.0 * (1.0 - vn * dtv / dy[nj - 1]) *
                                      (C(phi, i, nj) - C(phi, i, nj - 1)) / (YC(nj) - YC(nj - 1));
        else
            lst = C(phi, i, nj) - dy[nj - 2] / 2.0 * (1.0 + vn * dtv / dy[nj - 2]) *
                                      (C(phi, i, nj) - C(phi, i, nj - 1)) / (YC(nj) - YC(nj - 1));
        flux = lst * vn * dtv;
        C(temls, i, nj) = C(temls, i, nj) - flux / dy[nj - 1];
    }
}

/* Boundary_Condition, Clsvof_Mod.f90:961-973 */
void orc_boundary_copy(int ni, int nj, double* a)
{
#define A(i, j) a[((i)-1) + (size_t)((j)-1) * ni]
    for (int i = 1; i <= ni; i++) {
        A(i, 1) = A(i, 2);
        A(i, nj) = A(i, nj - 1);
    }
    for (int j = 1; j <= nj; j++) {
        A(1, j) = A(2, j);
        A(ni, j) = A(ni - 1, j);
    }
#undef A
}

/* isinsiderect, Clsvof_Mod.f90:1921-1955 (Heron areas of the four
 * triangles formed with the cell's edges). */
static int isinsiderect(double xp, double yp, double del_x, double del_y)
{
    double x1 = -0.5 * del_x, y1 = 0.5 * del_y;
    double x2 = 0.5 * del_x, y2 = 0.5 * del_y;
    double x3 = 0.5 * del_x, y3 = -0.5 * del_y;
    double x4 = -0.5 * del_x, y4 = -0.5 * del_y;
    double b1 = sqrt((x1 - xp) * (x1 - xp) + (y1 - yp) * (y1 - yp));
    double b2 = sqrt((x2 - xp) * (x2 - xp) + (y2 - yp) * (y2 - yp));
    double b3 = sqrt((x3 - xp) * (x3 - xp) + (y3 - yp) * (y3 - yp));
    double b4 = sqrt((x4 - xp) * (x4 - xp) + (y4 - yp) * (y4 - yp));
    double u1 = 0.5 * (del_x + b1 + b2);
    double u2 = 0.5 * (del_y + b2 + b3);
    double u3 = 0.5 * (del_x + b3 + b4);
    double u4 = 0.5 * (del_y + b4 + b1);
    double area = del_x * del_y;
    double area1 = sqrt(u1 * (u1 - del_x) * (u1 - b1) * (u1 - b2));
    double area2 = sqrt(u2 * (u2 - del_y) * (u2 - b2) * (u2 - b3));
    double area3 = sqrt(u3 * (u3 - del_x) * (u3 - b3) * (u3 - b4));
    double area4 = sqrt(u4 * (u4 - del_y) * (u4 - b4) * (u4 - b1));
    double areasum = area1 + area2 + area3 + area4;
    return fabs(areasum - area) < 1.0e-10;
}

/* Redistance, Clsvof_Mod.f90:1525-1708, in the reference's scatter order. */
void orc_redistance(const orc_cls_geom* g, const double* vfl, const double* nxx1, const double* nyy1,
                    const double* diss1, double* phi)
{
    int ni = g->ni, nj = g->nj;
    const double* vfls = g->vofS;
    const double* nxs = g->nxS;
    const double* nys = g->nyS;
    const double* phis = g->phiS;
    size_t n = (size_t)ni * nj;
    int* fix = (int*)calloc(n, sizeof(int));
    double* phiaux = (double*)malloc(n * sizeof(double));
    for (size_t k = 0; k < n; k++) phiaux[k] = 1.0e4;
    const double tol = 1.0e-24;
    for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) {
            if (!((0.0 + VOFEPS) < C(vfl, i, j) && C(vfl, i, j) < (1.0 - VOFEPS - C(vfls, i, j)))) continue;
            C(phiaux, i, j) = C(diss1, i, j);
            C(fix, i, j) = 1;
            double dxs = g->dx[i - 1], dys = g->dy[j - 1];
            double sxn = C(nxs, i, j), syn = C(nys, i, j), sph = C(phis, i, j);
            double n1 = C(nxx1, i, j), n2 = C(nyy1, i, j), d1 = C(diss1, i, j);
            for (int ii = -BAND_WIDTH; ii <= BAND_WIDTH; ii++) {
                for (int jj = -BAND_WIDTH; jj <= BAND_WIDTH; jj++) {
                    int ti = i + ii, tj = j + jj;
                    if (!(ti >= 1 && ti <= ni && tj >= 1 && tj <= nj)) continue;
                    double tv = C(vfl, ti, tj), ts = C(vfls, ti, tj);
                    double* pa = &C(phiaux, ti, tj);
                    if (tv < 0.0 + VOFEPS || tv >= 1.0 - VOFEPS - ts) {
                        C(fix, ti, tj) = 1;
                        int l = ii < -1 ? -1 : (ii > 1 ? 1 : ii);
                        int m = jj < -1 ? -1 : (jj > 1 ? 1 : jj);
                        double xv = dxs * (double)l / 2.0;
                        double yv = dys * (double)m / 2.0;
                        double dvs = sxn * xv + syn * yv + sph;
                        if (dvs < 0.0) {
                            if (fabs(yv) < VOFEPS) {
                                yv = dys / 2.0;
                                dvs = sxn * xv + syn * yv + sph;
                                if (dvs < 0.0) {
                                    yv = -dys / 2.0;
                                    dvs = sxn * xv + syn * yv + sph;
                                    if (dvs < 0.0) {
                                        xv = dxs * (0.0 - (double)l) / 2.0;
                                        dvs = sxn * xv + syn * yv + sph;
                                        if (dvs < 0.0) yv = dys / 2.0;
                                    }
                                }
                            } else if (fabs(xv) < VOFEPS) {
                                xv = dxs / 2.0;
                                dvs = sxn * xv + syn * yv + sph;
                                if (dvs < 0.0) {
                                    xv = -dxs / 2.0;
                                    dvs = sxn * xv + syn * yv + sph;
                                    if (dvs < 0.0) {
                                        yv = dys * (0.0 - (double)m) / 2.0;
                                        dvs = sxn * xv + syn * yv + sph;
                                        if (dvs < 0.0) xv = dxs / 2.0;
                                    }
                                }
                            } else {
                                yv = dys * (0.0 - (double)m) / 2.0;
                                dvs = sxn * xv + syn * yv + sph;
                                if (dvs < 0.0) {
                                    xv = dxs * (0.0 - (double)l) / 2.0;
                                    yv = dys * (double)m / 2.0;
                                    dvs = sxn * xv + syn * yv + sph;
                                    if (dvs < 0.0) yv = dys * (0.0 - (double)m) / 2.0;
                                }
                            }
                        }
                        double ddx = XC(ti) - XC(i), ddy = YC(tj) - YC(j);
                        double dv = n1 * xv + n2 * yv + d1;
                        if (dv * sgn(1.0, C(phi, ti, tj)) <= 0.0) {
                            double ax = ddx - xv, ay = ddy - yv;
                            double deuc = sqrt(ax * ax + ay * ay);
                            if (ts < EPSI) {
                                *pa = sgn(1.0, 0.5 - tv) * fmin(deuc, fabs(*pa));
                            } else {
                                if (fabs(deuc) < fabs(*pa) && tv >= 1.0 - VOFEPS - ts) *pa = -fabs(deuc);
                                if (fabs(deuc) < fabs(*pa) && tv < VOFEPS) *pa = fabs(deuc);
                            }
                        } else {
                            double dij1 = n1 * ddx + n2 * ddy + d1;
                            double xp = ddx - dij1 * n1;
                            double yp = ddy - dij1 * n2;
                            int pointp = isinsiderect(xp, yp, dxs, dys);
                            if (pointp && ts < EPSI) {
                                double ax = ddx - xp, ay = ddy - yp;
                                double deuc = sqrt(ax * ax + ay * ay);
                                *pa = sgn(1.0, 0.5 - tv) * fmin(deuc, fabs(*pa));
                            } else if (pointp && ts >= EPSI && ts < 1.0 - EPSI) {
                                double ax = ddx - xp, ay = ddy - yp;
                                double deuc = sqrt(ax * ax + ay * ay);
                                if (fabs(deuc) < fabs(*pa) && tv >= 1.0 - VOFEPS - ts) *pa = -fabs(deuc);
                                if (fabs(deuc) < fabs(*pa) && tv < VOFEPS) *pa = fabs(deuc);
                            } else if (ts >= 1.0 - EPSI) {
                                if (C(vfls, i, j) < EPSI) {
                                    double ax = ddx - xp, ay = ddy - yp;
                                    double deuc = sqrt(ax * ax + ay * ay);
                                    if (fabs(deuc) < fabs(*pa)) *pa = sgn(1.0, dij1) * deuc;
                                }
                            } else {
                                double xoff = fmax(fabs(xp) - 0.5 * dxs, 0.0);
                                double yoff = fmax(fabs(yp) - 0.5 * dys, 0.0);
                                double xfc = sgn(1.0, xp) * 0.5 * dxs;
                                double yfc = sgn(1.0, yp) * 0.5 * dys;
                                double xs, ys;
                                if (xoff * fabs(n1 + tol) >= yoff * fabs(n2 + tol)) {
                                    xs = xfc;
                                    ys = (d1 + n1 * xs) / (-n2 + tol);
                                } else {
                                    ys = yfc;
                                    xs = (d1 + n2 * ys) / (-n1 + tol);
                                }
                                double ax = ddx - xs, ay = ddy - ys;
                                double deuc = sqrt(ax * ax + ay * ay);
                                double dvs2 = sxn * xs + syn * ys + sph;
                                if (dvs2 >= 0.0) {
                                    if (ts < EPSI) {
                                        *pa = sgn(1.0, 0.5 - tv) * fmin(deuc, fabs(*pa));
                                    } else {
                                        if (fabs(deuc) < fabs(*pa) && tv >= 1.0 - VOFEPS - ts) *pa = -fabs(deuc);
                                        if (fabs(deuc) < fabs(*pa) && tv < VOFEPS) *pa = fabs(deuc);
                                    }
                                }
                            }
                        }
                    }
                    if (*pa > 9.999e3) C(fix, ti, tj) = 0;
                }
            }
        }
    }
    for (int i = 1; i <= ni; i++) {
        for (int j = 1; j <= nj; j++) {
            if (C(fix, i, j) == 1) {
                C(phi, i, j) = C(phiaux, i, j);
            } else {
                double dx = g->dx[i - 1], dy = g->dy[j - 1];
                if (C(vfl, i, j) <= VOFEPS)
                    C(phi, i, j) = (double)(BAND_WIDTH + 1) * sqrt(dx * dx + dy * dy);
                else if (C(vfl, i, j) >= (1.0 - VOFEPS - C(vfls, i, j)))
                    C(phi, i, j) = -(double)(BAND_WIDTH + 1) * sqrt(dx * dx + dy * dy);
            }
        }
    }
    free(fix);
    free(phiaux);
}

/* Depthw + etau of X_Sweep :1057-1059 for sub-step k */
double orc_inflow_level(const orc_cls_geom* g, const orc_wave* w, double time, int k, double dtv)
{
    double xwu = g->xc[0] - g->dx[0] / 2.0;
    double etau = w->amp * sin(w->kw * xwu - w->omega * (time + (double)k * dtv));
    return w->depth + etau;
}

/* Coupled_LS_VOF, Clsvof_Mod.f90:730-959, for a solid at rest (the
 * ObjectMovement / SolidVolumeFraction calls of :935-936 then reproduce the
 * geometry they are given).  vof, phi in/out; nx, ny out (PCell%nx/ny). */
void orc_clsvof_advect(const orc_cls_geom* g, double* vfl, double* phi, double* nx_out, double* ny_out,
                       const double* u, const double* v, double dt, double time, int nv, const orc_wave* w)
{
    int ni = g->ni, nj = g->nj;
    size_t n = (size_t)ni * nj;
    double* ue = (double*)malloc((size_t)(ni + 1) * nj * sizeof(double));
    double* ve = (double*)malloc((size_t)ni * (nj + 1) * sizeof(double));
    double* nx = (double*)malloc(n * sizeof(double));
    double* ny = (double*)malloc(n * sizeof(double));
    double* dis = (double*)malloc(n * sizeof(double));
    double* tvx = (double*)malloc(n * sizeof(double));
    double* tvy = (double*)malloc(n * sizeof(double));
    double* tlx = (double*)malloc(n * sizeof(double));
    double* tly = (double*)malloc(n * sizeof(double));
    orc_masked_velocities(g, u, v, ue, ve);
    double dtv = dt / (double)nv;
    const double* vS = g->vofS;
    for (int k = 1; k <= nv; k++) {
        double level = orc_inflow_level(g, w, time, k, dtv);
        if (k % 2 == 1) {
            memcpy(tvx, vfl, n * sizeof(double));
            memcpy(tlx, phi, n * sizeof(double));
            orc_reconstruct(g, vfl, phi, nx, ny, dis);
            orc_x_sweep(g, vfl, phi, nx, ny, dis, ue, dtv, level, tvx, tlx);
            for (int i = 1; i <= ni; i++)
                for (int j = 1; j <= nj; j++) {
                    double den = (1.0 - C(vS, i, j) + TOLDENO) -
                                 dtv / g->dx[i - 1] * (UE(i, j) * GH(g->aE, i, j) - UE(i - 1, j) * GH(g->aW, i, j));
                    C(tvx, i, j) = C(tvx, i, j) / den;
                    C(tlx, i, j) = C(tlx, i, j) / den;
                    C(vfl, i, j) = C(tvx, i, j) * (1.0 - C(vS, i, j));
                    C(phi, i, j) = C(tlx, i, j);
                    if (C(vfl, i, j) >= 1.0 - VOFEPS - C(vS, i, j)) C(vfl, i, j) = 1.0 - C(vS, i, j);
                    if (C(vfl, i, j) < VOFEPS) C(vfl, i, j) = 0.0;
                }
            orc_boundary_copy(ni, nj, vfl);
            orc_boundary_copy(ni, nj, phi);
            memcpy(tvy, vfl, n * sizeof(double));
            memcpy(tly, phi, n * sizeof(double));
            orc_reconstruct(g, vfl, phi, nx, ny, dis);
            orc_y_sweep(g, vfl, phi, nx, ny, dis, ve, dtv, tvy, tly);
            for (int i = 1; i <= ni; i++)
                for (int j = 1; j <= nj; j++) {
                    double dv = VE(i, j) * GH(g->aN, i, j) - VE(i, j - 1) * GH(g->aS, i, j);
                    C(tvy, i, j) = C(tvy, i, j) + dtv / g->dy[j - 1] * C(tvx, i, j) * dv;
                    C(tly, i, j) = C(tly, i, j) + dtv / g->dy[j - 1] * C(tlx, i, j) * dv;
                    C(vfl, i, j) = C(tvy, i, j);
                    if (C(vS, i, j) < 1.0 - EPSI) C(phi, i, j) = C(tly, i, j) / (1.0 - C(vS, i, j));
                }
        } else {
            memcpy(tvy, vfl, n * sizeof(double));
            memcpy(tly, phi, n * sizeof(double));
            orc_reconstruct(g, vfl, phi, nx, ny, dis);
            orc_y_sweep(g, vfl, phi, nx, ny, dis, ve, dtv, tvy, tly);
            for (int i = 1; i <= ni; i++)
                for (int j = 1; j <= nj; j++) {
                    double den = (1.0 - C(vS, i, j) + TOLDENO) -
                                 dtv / g->dy[j - 1] * (VE(i, j) * GH(g->aN, i, j) - VE(i, j - 1) * GH(g->aS, i, j));
                    C(tvy, i, j) = C(tvy, i, j) / den;
                    C(tly, i, j) = C(tly, i, j) / den;
                    C(vfl, i, j) = C(tvy, i, j) * (1.0 - C(vS, i, j));
                    C(phi, i, j) = C(tly, i, j);
                    if (C(vfl, i, j) >= 1.0 - VOFEPS - C(vS, i, j)) C(vfl, i, j) = 1.0 - C(vS, i, j);
                    if (C(vfl, i, j) < VOFEPS) C(vfl, i, j) = 0.0;
                }
            orc_boundary_copy(ni, nj, vfl);
            orc_boundary_copy(ni, nj, phi);
            memcpy(tvx, vfl, n * sizeof(double));
            memcpy(tlx, phi, n * sizeof(double));
            orc_reconstruct(g, vfl, phi, nx, ny, dis);
            orc_x_sweep(g, vfl, phi, nx, ny, dis, ue, dtv, level, tvx, tlx);
            for (int i = 1; i <= ni; i++)
                for (int j = 1; j <= nj; j++) {
                    double du = UE(i, j) * GH(g->aE, i, j) - UE(i - 1, j) * GH(g->aW, i, j);
                    C(tvx, i, j) = C(tvx, i, j) + dtv / g->dx[i - 1] * C(tvy, i, j) * du;
                    C(tlx, i, j) = C(tlx, i, j) + dtv / g->dx[i - 1] * C(tly, i, j) * du;
                    C(vfl, i, j) = C(tvx, i, j);
                    if (C(vS, i, j) < 1.0 - EPSI) C(phi, i, j) = C(tlx, i, j) / (1.0 - C(vS, i, j));
                }
        }
        orc_boundary_copy(ni, nj, vfl);
        orc_boundary_copy(ni, nj, phi);
        orc_reconstruct(g, vfl, phi, nx, ny, dis);
        memcpy(nx_out, nx, n * sizeof(double));
        memcpy(ny_out, ny, n * sizeof(double));
        orc_redistance(g, vfl, nx, ny, dis, phi);
    }
    free(ue); free(ve); free(nx); free(ny); free(dis);
    free(tvx); free(tvy); free(tlx); free(tly);
}

/* LagrangePoly, Clsvof_Mod.f90:712-726, at PosTar = 0 */
static double lagrange4(const double posi[4], const double valu[4])
{
    double val = 0.0;
    for (int i = 0; i < 4; i++) {
        double mult = valu[i];
        for (int j = 0; j < 4; j++)
            if (j != i) mult = mult * (0.0 - posi[j]) / (posi[i] - posi[j]);
        val = val + mult;
    }
    return val;
}

/* Liquid fraction of the half of P cell (i,j) on the U side (which=0: the
 * east half when `right` is 0, the west half of cell i+1 when 1) or the V
 * side, with the state branches of Initial_ClsVofUV :448-536 / :591-672.
 * Returns 0 when no branch assigns it (the reference keeps the previous
 * value). */
static int half_volf(const orc_cls_geom* g, const double* vof, const double* phi, const double* nx,
                     const double* ny, int which, int right, int i, int j, double phisT, double* out)
{
    double vf = C(vof, i, j), vs = C(g->vofS, i, j);
    double dx = g->dx[i - 1], dy = g->dy[j - 1];
    double ph;
    if (which == 0)
        ph = right ? C(phi, i, j) - dx / 4.0 * C(nx, i, j) : C(phi, i, j) + dx / 4.0 * C(nx, i, j);
    else
        ph = right ? C(phi, i, j) - dy / 4.0 * C(ny, i, j) : C(phi, i, j) + dy / 4.0 * C(ny, i, j);
    double hx = which == 0 ? dx / 2.0 : dx, hy = which == 0 ? dy : dy / 2.0;
    int set = 0;
    double v = 0.0, vols = 0.0;
    if (vf >= 1.0 - VOFEPS && vs < EPSI) { v = 1.0; set = 1; }
    if (vf < 1.0 - VOFEPS && vf >= VOFEPS && vs < EPSI) { v = orc_frac(C(nx, i, j), C(ny, i, j), ph, hx, hy); set = 1; }
    if (vf < VOFEPS && vs < EPSI) { v = 0.0; set = 1; }
    if (vf < VOFEPS && vs >= 1.0 - EPSI) { v = 0.0; set = 1; }
    if (vf < VOFEPS && vs < (which == 0 ? 1.0 - VOFEPS : 1.0 - EPSI) && vs >= EPSI) { v = 0.0; set = 1; }
    if (vf >= VOFEPS && vf < 1.0 - VOFEPS && vs >= EPSI && vs < 1.0 - EPSI) {
        orc_cellgeocal(C(g->nxS, i, j), C(g->nyS, i, j), phisT, C(nx, i, j), C(ny, i, j), ph, hx, hy, &vols, &v);
        set = 1;
    }
    if (set) *out = v;
    return set;
}

/* Initial_ClsVofUV, Clsvof_Mod.f90:402-710, for U (which=0) or V (which=1)
 * cells: level set and normal interpolated from the P cells (4-point
 * Lagrange, 2-point average near the edges), liquid fraction from the two
 * half P cells.  vofST, phiST: the staggered cells' solid fraction and solid
 * level set.  Outputs are cell arrays of the staggered grid. */
void orc_interface_to_uv(const orc_cls_geom* g, int which, const double* vof, const double* phi, const double* nx,
                         const double* ny, const double* vofST, const double* phiST, double* tvof, double* tphi,
                         double* tnx, double* tny)
{
    int ni = g->ni, nj = g->nj;
    double volfl = 0.0, volfr = 0.0; /* kept across cells, as in the reference */
    int nout = which == 0 ? nj : ni, nin = which == 0 ? ni - 1 : nj - 1;
    for (int o = 1; o <= nout; o++) {
        for (int k = 1; k <= nin; k++) {
            int i = which == 0 ? k : o, j = which == 0 ? o : k;
            int ip = which == 0 ? i + 1 : i, jp = which == 0 ? j : j + 1;
            int lagr = which == 0 ? (i >= 3 && i <= ni - 2) : (j >= 3 && j <= nj - 2);
            double p, a, b;
            if (lagr) {
                double posi[4], v1[4], v2[4], v3[4];
                for (int m = 0; m < 4; m++) {
                    int ii = which == 0 ? i - 1 + m : i, jj = which == 0 ? j : j - 1 + m;
                    if (which == 0)
                        posi[m] = (XC(i) + 0.5 * g->dx[i - 1]) - XC(ii);
                    else
                        posi[m] = (YC(j) + 0.5 * g->dy[j - 1]) - YC(jj);
                    v1[m] = C(phi, ii, jj);
                    v2[m] = C(nx, ii, jj);
                    v3[m] = C(ny, ii, jj);
                }
                p = lagrange4(posi, v1);
                a = lagrange4(posi, v2);
                b = lagrange4(posi, v3);
            } else {
                p = 0.5 * (C(phi, ip, jp) + C(phi, i, j));
                a = 0.5 * (C(nx, ip, jp) + C(nx, i, j));
                b = 0.5 * (C(ny, ip, jp) + C(ny, i, j));
            }
            double temn = which == 0 ? sqrt(a * a + b * b) + 1.0e-14 : sqrt(a * a + b * b + 1.0e-30);
            C(tphi, i, j) = p;
            C(tnx, i, j) = a / temn;
            C(tny, i, j) = b / temn;
            double phisT = C(phiST, i, j);
            half_volf(g, vof, phi, nx, ny, which, 0, i, j, phisT, &volfl);
            half_volf(g, vof, phi, nx, ny, which, 1, ip, jp, phisT, &volfr);
            double t = 0.5 * (volfl + volfr);
            if (t < VOFEPS) t = 0.0;
            if (t > 1.0 - VOFEPS - C(vofST, i, j)) t = 1.0 - C(vofST, i, j);
            int im = which == 0 ? i - 1 : i, jm = which == 0 ? j : j - 1;
            if ((which == 0 ? i > 1 : j > 1) && C(vof, im, jm) >= 1.0 - VOFEPS - C(g->vofS, im, jm) &&
                C(vof, ip, jp) >= 1.0 - VOFEPS - C(g->vofS, ip, jp))
                t = 1.0 - C(vofST, i, j);
            C(tvof, i, j) = t;
        }
        /* the last staggered cell copies its neighbour (:555-558 / :700-703) */
        int i = which == 0 ? ni : o, j = which == 0 ? o : nj;
        int im = which == 0 ? ni - 1 : o, jm = which == 0 ? o : nj - 1;
        C(tphi, i, j) = C(tphi, im, jm);
        C(tnx, i, j) = C(tnx, im, jm);
        C(tny, i, j) = C(tny, im, jm);
        C(tvof, i, j) = C(tvof, im, jm);
    }
}
