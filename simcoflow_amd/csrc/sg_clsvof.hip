// Coupled level-set / volume-of-fluid transport on gfx950 (Clsvof_Mod.f90).
//
// One call advances vof and phi over dt in nsub sub-steps (Coupled_LS_VOF,
// Clsvof_Mod.f90:730-959).  Each sub-step runs these launches:
//
//   k_recon      PLIC reconstruction (Interface_Reconstruct :1250-1321):
//                classify the cells and list the interface ones,
//   k_recon_ifc  then their normals and plane constants, packed one per lane
//   k_dis_fix    cells whose plane constant the reference leaves stale
//   k_flux_pre   PLIC face fluxes of the listed cells, packed one per lane
//   k_sweep      first sweep, implicit divergence (:798-823 / :862-888)
//   k_bc         boundary copy (Boundary_Condition :961-973)
//   k_recon, k_recon_ifc, k_dis_fix, k_flux_pre
//   k_sweep      second sweep, explicit divergence (:834-858 / :898-929)
//   k_bc, k_recon (+ interface bit mask and source list), k_recon_ifc, k_dis_fix
//   k_rd_scatter, k_rd_dilate, k_redistance_far, k_redistance_band: level
//                set from the PLIC segments (Redistance :1525-1708)
//
// The reference's sweeps walk each row in order and update cells in place;
// every face flux they use depends only on data from before the sweep, so a
// sweep here is one thread per cell that recomputes the fluxes of its two
// faces and applies them in the reference's order.  Redistance scatters
// from interface cells into a 9x9 window.  For two- and three-phase targets
// the result is a minimum over the visits, taken here with one lane per
// (source, window offset) pair and an integer atomic minimum on the
// distance's bit pattern; full-solid targets, whose sign depends on the
// order of equal distances, replay their visits in the reference's order.
// Both give the reference's result bit for bit.
//
// This file is compiled with FP contraction off: every sum and product
// rounds as the reference writes it.
#include "sg_internal.hpp"
#include "sg_poly.hpp"

namespace sg {

namespace {

constexpr double VOFEPS = 1.0e-14;   // Clsvof_Mod.f90:10
constexpr double TOLDENO = 1.0e-24;  // Clsvof_Mod.f90:10
constexpr int BAND = 4;              // Clsvof_Mod.f90:9 band_width

struct ClsGeo {
    int ni, nj;
    const double *dx, *dy, *xc, *yc;
    const double *vofS, *nxS, *nyS, *phiS;
    const double *aE, *aW, *aN, *aS;  // ghosted
    const int *uflag, *vflag;
    const double *u, *v;  // ghosted TVar%u, TVar%v
};

// 1-based cell index (the reference's (i,j)) -> cell array offset
__device__ __forceinline__ size_t cix(const ClsGeo& G, int i, int j) {
    return (size_t)(i - 1) + (size_t)(j - 1) * (size_t)G.ni;
}
__device__ __forceinline__ size_t gix(const ClsGeo& G, int i, int j) {
    return (size_t)i + (size_t)j * (size_t)(G.ni + 2);
}
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

// ---- PLIC geometry --------------------------------------------------------

// frac (:1371-1404): liquid fraction of a dx*dy box below |nx|x+|ny|y+d=0
__device__ double plic_frac(double nx, double ny, double d, double dx, double dy) {
    const double tnx = fabs(nx), tny = fabs(ny);
    if (d + 0.5 * (tnx * dx + tny * dy) <= 0.0) return 1.0;
    if (d - 0.5 * (tnx * dx + tny * dy) >= 0.0) return 0.0;
    if (tnx <= 1.0e-20) return 0.5 - d / dy;
    if (tny <= 1.0e-20) return 0.5 - d / dx;
    const double xx = (0.5 * tny * dy - d) / dx / tnx + 0.5;
    const double yy = (0.5 * tnx * dx - d) / dy / tny + 0.5;
    const double area = 0.5 * xx * yy;
    double top = 0.0, right = 0.0;
    if (yy > 1.0) {
        const double t = (yy - 1.0) / yy;
        top = t * t;
    }
    if (xx > 1.0) {
        const double t = (xx - 1.0) / xx;
        right = t * t;
    }
    return area * (1.0 - top - right);
}

// distance (:1323-1369): plane constant for volume fraction vf
__device__ double plic_distance(double vf, double nxx, double nyy, double dx, double dy) {
    const double nx = fabs(nxx), ny = fabs(nyy);
    if (nx < 1.0e-20) return (0.5 - vf) * dy;
    if (ny < 1.0e-20) return (0.5 - vf) * dx;
    const double dll = 0.5 * (nx * dx + ny * dy);
    double dul = 0.5 * (nx * dx - ny * dy);
    double dlr = -dul;
    const double dur = -dll;
    double vul = plic_frac(nx, ny, dul, dx, dy);
    double vlr = plic_frac(nx, ny, dlr, dx, dy);
    if (vf <= vul && vf <= vlr) {
        if (vul > vlr) {
            double t = vul; vul = vlr; vlr = t;
            t = dul; dul = dlr; dlr = t;
        }
        const double e = dul - dll;
        return dll - sqrt((e * e) * vf / vul);
    }
    if (vf >= vul && vf >= vlr) {
        if (vlr < vul) {
            double t = vul; vul = vlr; vlr = t;
            t = dul; dul = dlr; dlr = t;
        }
        const double e = dlr - dur;
        return dur + sqrt(e * e * (1.0 - vf) / (1.0 - vlr));
    }
    if (fabs(vlr - vul) < 1.0e-20) return 0.5 * (dul + dlr);
    const double th = (vf - vul) / (vlr - vul);
    return dlr * th + dul * (1.0 - th);
}

// Fluid part of a cell cut by the solid line, counter-clockwise from the SE
// corner (CellGeoCal :2170-2219, DistanceFluidCalculate :1966-2015): n <= 5
// vertices, x[n] repeats x[0], the rest are zero.  Every loop over the
// vertices is unrolled with a guard in place of a run-time bound, so each
// index is a compile-time constant and the polygon stays in registers.
struct Poly {
    double x[6], y[6];
    int n;
};

__device__ __forceinline__ void solid_poly(double nxs, double nys, double phis, double dx, double dy, Poly& P) {
#pragma unroll
    for (int k = 0; k < 6; ++k) P.x[k] = P.y[k] = 0.0;
    int n = 0;
    fluid_vertices(nxs, nys, phis, dx, dy, [&](double a, double b) {
        if (n < 5) {
#pragma unroll
            for (int s = 0; s < 5; ++s)
                if (s == n) {
                    P.x[s] = a;
                    P.y[s] = b;
                }
            ++n;
        }
    });
    const double x0 = P.x[0], y0 = P.y[0];
#pragma unroll
    for (int s = 1; s < 6; ++s)
        if (s == n) {
            P.x[s] = x0;
            P.y[s] = y0;
        }
    P.n = n;
}

// CellGeoCal (:2159-2253)
__device__ __forceinline__ void cell_geo(double nxs, double nys, double phis, double nxl, double nyl, double phil,
                                         double dx, double dy, double& vols, double& volf) {
    cut_cell_fractions(nxs, nys, phis, nxl, nyl, phil, dx, dy, vols, volf);
}

// LiqFuncVol (:2107-2155): volf less the fraction of the fluid polygon on
// the liquid side of the line through (xf, yf) with normal (nxf, nyf).  The
// reference takes the vertex count from the last non-zero vertex, nz - 1
// (the repeat of the first vertex makes that n), and closes the polygon at
// vertex nz - 1 with the first vertex's distance.  Inlined, so P stays in
// registers.
__device__ __forceinline__ double liq_func(double nxf, double nyf, double volf, double dx, double dy, const Poly& P, double xf,
                           double yf) {
    const double phil = -(nxf * xf + nyf * yf);
    // the reference counts vertices as the last non-zero node
    int nz = 0;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        if (fabs(P.x[k]) > 1.0e-14 || fabs(P.y[k]) > 1.0e-14) nz = k + 1;
    if (nz == 0) return volf;
    const int n = nz - 1;
    double pdt[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) pdt[k] = P.x[k] * nxf + P.y[k] * nyf + phil;
    ShoelaceClip c;
#pragma unroll
    for (int k = 0; k < 5; ++k)
        if (k < n) c.edge(P.x[k], P.y[k], pdt[k], P.x[k + 1], P.y[k + 1], k + 1 == n ? pdt[0] : pdt[k + 1]);
    return volf - fabs(c.close() / (dx * dy));
}

// DistanceFluidCalculate + zriddr (:1957-2105).  Returns false when the
// reference leaves the plane constant unassigned.
__device__ bool distance_fluid(double nxs, double nys, double phis, double nxf, double nyf, double volf, double dx,
                               double dy, double& phif) {
    Poly P;
    solid_poly(nxs, nys, phis, dx, dy, P);
    const double diag = sqrt(dx * dx + dy * dy);
    double maxd = -2.0 * diag, mind = 2.0 * diag;
    double hx = 0.0, hy = 0.0, lx = 0.0, ly = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        if (k >= P.n) break;
        const double t = nxf * P.x[k] + nyf * P.y[k];
        if (t > maxd) {
            maxd = t;
            hx = P.x[k];
            hy = P.y[k];
        }
        if (t < mind) {
            mind = t;
            lx = P.x[k];
            ly = P.y[k];
        }
    }
    hx = hx + hx * nxf * EPSI;
    hy = hy + hy * nyf * EPSI;
    lx = lx - lx * nxf * EPSI;
    ly = ly - ly * nyf * EPSI;
    // Ridders' method on the segment (lx,ly) -> (hx,hy), tol 1e-14, 1000 its
    const double tol = 1.0e-14;
    double fl = liq_func(nxf, nyf, volf, dx, dy, P, lx, ly);
    double fh = liq_func(nxf, nyf, volf, dx, dy, P, hx, hy);
    if ((fl > tol && fh < -tol) || (fl < -tol && fh > tol)) {
        double xlo = lx, ylo = ly, xhi = hx, yhi = hy;
        double ax = -1.11e-30, ay = -1.11e-30;
        for (int it = 0; it < 1000; ++it) {
            const double mx = 0.5 * (xlo + xhi), my = 0.5 * (ylo + yhi);
            const double fm = liq_func(nxf, nyf, volf, dx, dy, P, mx, my);
            const double s = sqrt(fm * fm - fl * fh);
            const double sg = fsign(1.0, fl - fh);
            const double nxw = mx + (mx - xlo) * sg * fm / s;
            const double nyw = my + (my - ylo) * sg * fm / s;
            const double e0 = nxw - ax, e1 = nyw - ay;
            if (sqrt(e0 * e0 + e1 * e1) <= tol) {
                phif = -(nxf * ax + nyf * ay);
                return true;
            }
            ax = nxw;
            ay = nyw;
            const double fn = liq_func(nxf, nyf, volf, dx, dy, P, ax, ay);
            if (fabs(fn) <= tol) {
                phif = -(nxf * ax + nyf * ay);
                return true;
            }
            if (fsign(fm, fn) != fm) {
                xlo = mx; ylo = my; fl = fm;
                xhi = ax; yhi = ay; fh = fn;
            } else if (fsign(fl, fn) != fl) {
                xhi = ax; yhi = ay; fh = fn;
            } else if (fsign(fh, fn) != fh) {
                xlo = ax; ylo = ay; fl = fn;
            }
            const double h0 = xhi - xlo, h1 = yhi - ylo;
            if (sqrt(h0 * h0 + h1 * h1) <= tol) {
                phif = -(nxf * ax + nyf * ay);
                return true;
            }
        }
        phif = -(nxf * ax + nyf * ay);
        return true;
    }
    if (fl < tol && fl > -tol) {
        phif = -(nxf * lx + nyf * ly);
        return true;
    }
    if (fh < tol && fh > -tol) {
        phif = -(nxf * hx + nyf * hy);
        return true;
    }
    return false;
}

// ---- normals ----------------------------------------------------------------

struct PhiView {
    const ClsGeo& G;
    const double* phi;
    __device__ double operator()(int i, int j) const { return phi[cix(G, i, j)]; }
    __device__ double X(int i) const { return G.xc[i - 1]; }
    __device__ double Y(int j) const { return G.yc[j - 1]; }
};

__device__ __forceinline__ double qual(double vx, double vy) { return fabs(1.0 - sqrt(vx * vx + vy * vy)); }

// The 8 candidate directions of Coupled_LS_VOF (:746-753) as the reference
// leaves them: Vset(1,2) is overwritten with +1 and Vset(2,2) never set, so
// direction 2 is (0,0) and drops out (its cosine is 0/0).
__device__ __forceinline__ void vset(int k, double& a, double& b) {
    const double A[8] = {0.0, 0.0, -1.0, -1.0, -1.0, 1.0, 1.0, 1.0};
    const double B[8] = {1.0, 0.0, 0.0, -1.0, 1.0, 0.0, -1.0, 1.0};
    a = A[k];
    b = B[k];
}

// one-sided pair chosen by the direction most parallel to grad phi
// (Normal_Vector_Irre :1768-1827 / :1855-1905)
__device__ void nv_directional(const PhiView& p, int i, int j, int imv0, double& nx, double& ny) {
    const double eta = 0.075, nuy = eta / 8.0;
    const double vx = (p(i + 1, j) - p(i - 1, j)) / (p.X(i + 1) - p.X(i - 1));
    const double vy = (p(i, j + 1) - p(i, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
    double maxv = 0.0;
    int imv = imv0;
    for (int k = 1; k <= 8; ++k) {
        double a, b;
        vset(k - 1, a, b);
        const double cp = fabs(a * vx + b * vy) / (sqrt(a * a + b * b) * sqrt(vx * vx + vy * vy));
        if (cp >= maxv) {
            maxv = cp;
            imv = k;
        }
    }
    double v1x, v1y;
    if (imv == 0) {  // zero gradient: outside Vset in the reference
        v1x = -1.0;
        v1y = 0.0;
    } else {
        double a, b;
        vset(imv - 1, a, b);
        v1x = -b;
        v1y = a;
    }
    double vx1, vx2, vy1, vy2;
    if (v1x == -1.0) {
        vx1 = (p(i, j) - p(i - 1, j)) / (p.X(i) - p.X(i - 1));
        vx2 = (p(i + 1, j) - p(i, j)) / (p.X(i + 1) - p.X(i));
    } else if (v1x == 0.0) {
        vx1 = (p(i + 1, j) - p(i - 1, j)) / (p.X(i + 1) - p.X(i - 1));
        vx2 = vx1;
    } else {
        vx1 = (p(i + 1, j) - p(i, j)) / (p.X(i + 1) - p.X(i));
        vx2 = (p(i, j) - p(i - 1, j)) / (p.X(i) - p.X(i - 1));
    }
    if (v1y == -1.0) {
        vy1 = (p(i, j) - p(i, j - 1)) / (p.Y(j) - p.Y(j - 1));
        vy2 = (p(i, j + 1) - p(i, j)) / (p.Y(j + 1) - p.Y(j));
    } else if (v1y == 0.0) {
        vy1 = (p(i, j + 1) - p(i, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
        vy2 = vy1;
    } else {
        vy1 = (p(i, j + 1) - p(i, j)) / (p.Y(j + 1) - p.Y(j));
        vy2 = (p(i, j) - p(i, j - 1)) / (p.Y(j) - p.Y(j - 1));
    }
    if (qual(vx1, vy1) < qual(vx2, vy2) + nuy) {
        nx = vx1;
        ny = vy1;
    } else {
        nx = vx2;
        ny = vy2;
    }
}

// Normal_Vector_Irre (:1710-1919), interior cells
__device__ void normal_irre(const ClsGeo& G, const double* phi, int i, int j, double& nx, double& ny) {
    const PhiView p{G, phi};
    const double eta = 0.075;
    const double vs = G.vofS[cix(G, i, j)];
    if (vs >= EPSI) {
        const double se = G.vofS[cix(G, i + 1, j)], sw = G.vofS[cix(G, i - 1, j)];
        const double sn = G.vofS[cix(G, i, j + 1)], ss = G.vofS[cix(G, i, j - 1)];
        if (se < 1.0 - EPSI && sw < 1.0 - EPSI)
            nx = (p(i + 1, j) - p(i - 1, j)) / (p.X(i + 1) - p.X(i - 1));
        else if (se >= 1.0 - EPSI)
            nx = (p(i, j) - p(i - 1, j)) / (p.X(i) - p.X(i - 1));
        else
            nx = (p(i + 1, j) - p(i, j)) / (p.X(i + 1) - p.X(i));
        if (sn < 1.0 - EPSI && ss < 1.0 - EPSI)
            ny = (p(i, j + 1) - p(i, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
        else if (sn > 1.0 - EPSI)
            ny = (p(i, j) - p(i, j - 1)) / (p.Y(j) - p.Y(j - 1));
        else
            ny = (p(i, j + 1) - p(i, j)) / (p.Y(j + 1) - p.Y(j));
        return;
    }
    const double cx = (p(i + 1, j) - p(i - 1, j)) / (p.X(i + 1) - p.X(i - 1));
    const double cy = (p(i, j + 1) - p(i, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
    const double q0 = qual(cx, cy);
    double vx, vy;
    vx = i >= 3 ? (p(i, j) - p(i - 2, j)) / (p.X(i) - p.X(i - 2)) : (p(i, j) - p(i - 1, j)) / (p.X(i) - p.X(i - 1));
    vy = (p(i - 1, j + 1) - p(i - 1, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
    const double qm = qual(vx, vy);
    vx = i <= G.ni - 2 ? (p(i + 2, j) - p(i, j)) / (p.X(i + 2) - p.X(i))
                       : (p(i + 1, j) - p(i, j)) / (p.X(i + 1) - p.X(i));
    vy = (p(i + 1, j + 1) - p(i + 1, j - 1)) / (p.Y(j + 1) - p.Y(j - 1));
    const double qp = qual(vx, vy);
    int sx;
    if (qm < eta && qp >= eta) sx = -1;
    else if (qm >= eta && qp < eta) sx = 1;
    else if (qm < eta && q0 < eta && qp < eta) sx = 0;
    else if (qm >= eta && q0 >= eta && qp >= eta) sx = 0;
    else {
        nv_directional(p, i, j, 0, nx, ny);
        return;
    }
    vy = j >= 3 ? (p(i, j) - p(i, j - 2)) / (p.Y(j) - p.Y(j - 2)) : (p(i, j) - p(i, j - 1)) / (p.Y(j) - p.Y(j - 1));
    vx = (p(i + 1, j - 1) - p(i - 1, j - 1)) / (p.X(i + 1) - p.X(i - 1));
    const double rm = qual(vx, vy);
    vx = (p(i + 1, j + 1) - p(i - 1, j + 1)) / (p.X(i + 1) - p.X(i - 1));
    vy = j <= G.nj - 2 ? (p(i, j + 2) - p(i, j)) / (p.Y(j + 2) - p.Y(j))
                       : (p(i, j + 1) - p(i, j)) / (p.Y(j + 1) - p.Y(j));
    const double rp = qual(vx, vy);
    int sy;
    if (rm < eta && rp >= eta) sy = -1;
    else if (rm >= eta && rp < eta) sy = 1;
    else if (rm < eta && q0 < eta && rp < eta) sy = 0;
    else if (rm >= eta && q0 >= eta && rp >= eta) sy = 0;
    else {
        nv_directional(p, i, j, 1, nx, ny);
        return;
    }
    if (sx == -1) nx = (p(i, j) - p(i - 1, j)) / (p.X(i) - p.X(i - 1));
    else if (sx == 1) nx = (p(i + 1, j) - p(i, j)) / (p.X(i + 1) - p.X(i));
    else nx = cx;
    if (sy == -1) ny = (p(i, j) - p(i, j - 1)) / (p.Y(j) - p.Y(j - 1));
    else if (sy == 1) ny = (p(i, j + 1) - p(i, j)) / (p.Y(j + 1) - p.Y(j));
    else ny = cy;
}

// interface class of Interface_Reconstruct (:1258-1266): 0 none, 1 two
// phases, 2 three phases
__device__ __forceinline__ int recon_class(double vf, double vs) {
    if (vf >= VOFEPS && vf < (1.0 - VOFEPS - vs)) {
        if (vs < EPSI) return 1;
        if (vs < 1.0 - EPSI) return 2;
    }
    return 0;
}

__device__ __forceinline__ void normalise(double& nx, double& ny) {
    const double t = sqrt(nx * nx + ny * ny);
    if (t < 1.0e-14) {
        nx = 0.0;
        ny = 1.0;
    } else {
        nx = nx / t;
        ny = ny / t;
    }
}

// one-sided differences the reference uses on the domain edge (:1272-1283)
__device__ __forceinline__ double edge_nx(const PhiView& p, int i, int j, int ni) {
    return i == 1 ? (p(2, j) - p(1, j)) / (p.X(2) - p.X(1)) : (p(ni, j) - p(ni - 1, j)) / (p.X(ni) - p.X(ni - 1));
}
__device__ __forceinline__ double edge_ny(const PhiView& p, int i, int j, int nj) {
    return j == 1 ? (p(i, 2) - p(i, 1)) / (p.Y(2) - p.Y(1)) : (p(i, nj) - p(i, nj - 1)) / (p.Y(nj) - p.Y(nj - 1));
}

__device__ __forceinline__ int cls_of(const ClsGeo& G, const double* vof, int i, int j) {
    const size_t c = cix(G, i, j);
    return recon_class(vof[c], G.vofS[c]);
}

// final normal of an interior cell
__device__ void final_interior(const ClsGeo& G, const double* vof, const double* phi, int i, int j, double& nx,
                               double& ny) {
    if (cls_of(G, vof, i, j)) {
        normal_irre(G, phi, i, j, nx, ny);
        normalise(nx, ny);
    } else {
        nx = 0.0;
        ny = 1.0;
    }
}

// final normal of a cell on the west or east edge (i = 1 or ni).  Away from
// the corners the reference's y component is the previous cell's (i,j-1)
// final value (a stale local, :1271-1283), so a run of interface cells up the
// edge is a chain; it is walked from the cell below its start.
__device__ void final_side(const ClsGeo& G, const double* vof, const double* phi, int i, int j, double& nx,
                           double& ny) {
    const PhiView p{G, phi};
    const int nj = G.nj;
    if (!cls_of(G, vof, i, j)) {
        nx = 0.0;
        ny = 1.0;
        return;
    }
    if (j == 1 || j == nj) {
        nx = edge_nx(p, i, j, G.ni);
        ny = edge_ny(p, i, j, nj);
        normalise(nx, ny);
        return;
    }
    int js = j - 1;
    while (js > 1 && cls_of(G, vof, i, js)) --js;
    double nyp;
    if (js == 1 && cls_of(G, vof, i, 1)) {
        double a = edge_nx(p, i, 1, G.ni), b = edge_ny(p, i, 1, nj);
        normalise(a, b);
        nyp = b;
    } else {
        nyp = 1.0;
    }
    for (int k = js + 1; k <= j; ++k) {
        double a = edge_nx(p, i, k, G.ni), b = nyp;
        normalise(a, b);
        nyp = b;
        if (k == j) {
            nx = a;
            ny = b;
        }
    }
}

// final normal of a cell on the south or north edge (1 < i < ni).  Its x
// component is the previous cell's in the reference's loop order: (i,nj-1)
// for the north edge, (i-1,nj) for the south edge.
__device__ void final_topbottom(const ClsGeo& G, const double* vof, const double* phi, int i, int j, double& nx,
                                double& ny) {
    const PhiView p{G, phi};
    const int ni = G.ni, nj = G.nj;
    if (!cls_of(G, vof, i, j)) {
        nx = 0.0;
        ny = 1.0;
        return;
    }
    double snx, t;
    if (j == nj) {
        final_interior(G, vof, phi, i, nj - 1, snx, t);
    } else {  // j == 1: previous cell is (i-1, nj)
        const int ip = i - 1;
        if (ip == 1) {
            final_side(G, vof, phi, 1, nj, snx, t);
        } else if (cls_of(G, vof, ip, nj)) {
            double a, b;
            final_interior(G, vof, phi, ip, nj - 1, a, b);
            b = edge_ny(p, ip, nj, nj);
            normalise(a, b);
            snx = a;
        } else {
            snx = 0.0;
        }
    }
    nx = snx;
    ny = edge_ny(p, i, j, nj);
    normalise(nx, ny);
}

// ---- kernels ------------------------------------------------------------------

// PLIC planes of the last reconstruction.  They are stored only for interface
// cells (class 1 or 2); every other cell has n = (0,1) and its plane constant
// follows from its volume fraction (Interface_Reconstruct :1307-1314), so
// readers rebuild those on the fly instead of reading them.
struct Planes {
    double* nx;
    double* ny;
    double* dis;
};

__device__ __forceinline__ double default_dis(double vf, double vs, double hy) {
    return 1.0 - vs > VOFEPS ? (0.5 - vf / (1.0 - vs)) * hy : (0.0 - vf) * hy;
}

__device__ __forceinline__ void plane_of(const ClsGeo& G, const Planes& P, size_t c, double vf, double hy, double& nx,
                                         double& ny, double& d) {
    const double vs = G.vofS[c];
    if (recon_class(vf, vs)) {
        nx = P.nx[c];
        ny = P.ny[c];
        d = P.dis[c];
    } else {
        nx = 0.0;
        ny = 1.0;
        d = default_dis(vf, vs, hy);
    }
}
// plane_of from values already loaded (cnx, cny, cdis: the stored plane)
__device__ __forceinline__ void plane_pick(double vf, double vs, double hy, double cnx, double cny, double cdis,
                                           double& nx, double& ny, double& d) {
    if (recon_class(vf, vs)) {
        nx = cnx;
        ny = cny;
        d = cdis;
    } else {
        nx = 0.0;
        ny = 1.0;
        d = default_dis(vf, vs, hy);
    }
}

// wavefront-aggregated append of `value` to list where pred holds (one
// atomic per wavefront); every lane of the wavefront must call it
__device__ __forceinline__ void append(bool pred, unsigned* count, int* list, int value) {
    const unsigned long long b = __ballot(pred);
    if (!b) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)b) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(b));
    base = __shfl(base, leader);
    if (pred) list[base + (unsigned)__popcll(b & ((1ull << lane) - 1ull))] = value;
}

// Interface_Reconstruct.  Interface cells get their normal and plane
// constant; three-phase cells whose Ridders search leaves the plane constant
// unassigned (the reference then keeps the previous cell's) are flagged in
// `stale` and listed for k_dis_fix.  full_out: also write n = (0,1) for the
// other cells (the PCell%nx/ny output).  With mask != null the kernel also
// writes the re-distancing source mask (vofeps < vof < 1-vofeps-vofS, one bit
// per cell, one 64-bit word per wavefront and row) and lists those cells.
// cnt[0]: sources listed, cnt[1]: stale cells listed, cnt[2]: interface
// cells listed.
//
// This scan is bound by its appends, not its bytes: every workgroup that
// lists a cell takes one atomic on a counter all workgroups share.  So a
// workgroup covers 64 x 64 cells, each thread RECON_R rows, and lists all of
// them with one scan and one atomic per list (the two per-thread counts
// packed in one word, low and high 16 bits: at most 4096 of each per
// workgroup).
constexpr int RECON_TY = 16;  // wavefronts per k_recon workgroup
constexpr int RECON_R = 4;    // rows per thread: a workgroup covers 64 x (RECON_TY * RECON_R) cells
__global__ void __launch_bounds__(64 * RECON_TY) k_recon(ClsGeo G, const double* __restrict__ vof, Planes P,
                                                         int full_out,
                                               unsigned long long* __restrict__ mask, int mw,
                                               int* __restrict__ srcs, int* __restrict__ ifc,
                                               unsigned* __restrict__ cnt) {
    __shared__ unsigned wsum[RECON_TY], wbase_i[RECON_TY], wbase_s[RECON_TY], gbase[2];
    const int lane = threadIdx.x, w = threadIdx.y;
    const int i = blockIdx.x * 64 + lane + 1;
    const int j0 = blockIdx.y * (RECON_TY * RECON_R) + w + 1;
    // every row's operands loaded before any test
    double vf[RECON_R], vs[RECON_R];
#pragma unroll
    for (int r = 0; r < RECON_R; ++r) {
        const int j = j0 + r * RECON_TY;
        const bool in = i <= G.ni && j <= G.nj;
        const size_t c = in ? cix(G, i, j) : 0;
        vf[r] = in ? vof[c] : 0.0;
        vs[r] = in ? G.vofS[c] : 0.0;
    }
    unsigned list_i = 0, list_s = 0;  // bit r: row r's cell goes to ifc / srcs
#pragma unroll
    for (int r = 0; r < RECON_R; ++r) {
        const int j = j0 + r * RECON_TY;
        const bool in = i <= G.ni && j <= G.nj;
        const bool src = in && (0.0 + VOFEPS) < vf[r] && vf[r] < (1.0 - VOFEPS - vs[r]);
        const bool itf = in && recon_class(vf[r], vs[r]) != 0;
        const bool heavy = in && !(vf[r] >= (1.0 - VOFEPS) || vf[r] <= VOFEPS);  // vof_flux's PLIC branch
        if (in && !itf && full_out) {
            const size_t c = cix(G, i, j);
            P.nx[c] = 0.0;
            P.ny[c] = 1.0;
        }
        if (itf || heavy) list_i |= 1u << r;
        if (mask) {
            // one wavefront covers 64 consecutive i of row j
            const unsigned long long bits = __ballot(src);
            if (lane == 0 && j <= G.nj) mask[(size_t)(j - 1) * mw + blockIdx.x] = bits;
            if (src) list_s |= 1u << r;
        }
    }
    // exclusive offsets of this thread's entries in the workgroup's share of
    // each list: an inclusive scan of the packed counts over the wavefront,
    // then over the wavefronts
    const unsigned own = (unsigned)__popc(list_i) | ((unsigned)__popc(list_s) << 16);
    unsigned incl = own;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned t = __shfl_up(incl, o, 64);
        if (lane >= o) incl += t;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (w == 0 && lane == 0) {
        unsigned si = 0, ss = 0;
        for (int k = 0; k < RECON_TY; ++k) {
            wbase_i[k] = si;
            wbase_s[k] = ss;
            si += wsum[k] & 0xFFFFu;
            ss += wsum[k] >> 16;
        }
        gbase[0] = si ? atomicAdd(cnt + 2, si) : 0u;
        gbase[1] = ss ? atomicAdd(cnt, ss) : 0u;
    }
    __syncthreads();
    const unsigned excl = incl - own;
    unsigned oi = gbase[0] + wbase_i[w] + (excl & 0xFFFFu);
    unsigned os = gbase[1] + wbase_s[w] + (excl >> 16);
#pragma unroll
    for (int r = 0; r < RECON_R; ++r) {
        const int c = (int)cix(G, i, j0 + r * RECON_TY);
        if (list_i >> r & 1u) ifc[oi++] = c;
        if (list_s >> r & 1u) srcs[os++] = c;
    }
}

// The interface cells listed by k_recon, packed one per lane: the normal and
// plane constant cost the same per cell wherever a cell lies, and packing
// keeps the wavefronts of bulk liquid and gas out of that work.
__global__ void __launch_bounds__(256) k_recon_ifc(ClsGeo G, const double* __restrict__ vof,
                                                   const double* __restrict__ phi, Planes P,
                                                   unsigned char* __restrict__ stale, const int* __restrict__ ifc,
                                                   const unsigned* __restrict__ nifc, int* __restrict__ stales,
                                                   unsigned* __restrict__ cnt) {
    const unsigned n = *nifc;
    for (unsigned base = blockIdx.x * blockDim.x; base < n; base += gridDim.x * blockDim.x) {
        const unsigned t = base + threadIdx.x;
        bool st = false;
        int c = 0;
        if (t < n) c = ifc[t];
        const double vf = t < n ? vof[c] : 0.0, vs = t < n ? G.vofS[c] : 0.0;
        const int cl = t < n ? recon_class(vf, vs) : 0;
        if (cl) {
            const int i = c % G.ni + 1, j = c / G.ni + 1;
            double nx, ny, d = 0.0;
            if (i > 1 && i < G.ni && j > 1 && j < G.nj)
                final_interior(G, vof, phi, i, j, nx, ny);
            else if (i == 1 || i == G.ni)
                final_side(G, vof, phi, i, j, nx, ny);
            else
                final_topbottom(G, vof, phi, i, j, nx, ny);
            const double hx = G.dx[i - 1], hy = G.dy[j - 1];
            if (cl == 1)
                d = plic_distance(vf, nx, ny, hx, hy);
            else
                st = !distance_fluid(G.nxS[c], G.nyS[c], G.phiS[c], nx, ny, vf, hx, hy, d);
            P.nx[c] = nx;
            P.ny[c] = ny;
            P.dis[c] = d;
            if (cl == 2) stale[c] = st ? 1 : 0;
        }
        append(st, cnt + 1, stales, c);
    }
}

// Cells whose plane constant the reference leaves stale take the value of
// the nearest earlier cell in its loop order (i outer, j inner) that has one.
__global__ void __launch_bounds__(256) k_dis_fix(ClsGeo G, const double* __restrict__ vof,
                                                 const unsigned char* __restrict__ stale,
                                                 const int* __restrict__ list, const unsigned* __restrict__ cnt,
                                                 double* __restrict__ dis) {
    const unsigned n = cnt[1];
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int c0 = list[t];
        int pi = c0 % G.ni + 1, pj = c0 / G.ni + 1;
        double d = 0.0;
        for (;;) {
            if (pj > 1) {
                --pj;
            } else if (pi > 1) {
                --pi;
                pj = G.nj;
            } else {
                break;  // before the first cell: the reference's uninitialised local
            }
            const size_t p = cix(G, pi, pj);
            const double vf = vof[p], vs = G.vofS[p];
            const int cl = recon_class(vf, vs);
            if (cl == 2 && stale[p]) continue;
            d = cl ? dis[p] : default_dis(vf, vs, G.dy[pj - 1]);
            break;
        }
        dis[c0] = d;
    }
}

// Boundary_Condition (:961-973) on two arrays: every ring cell takes its
// value from an interior cell, so the copy is race-free in place.
__global__ void k_bc(int ni, int nj, double* __restrict__ a, double* __restrict__ b) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, si, sj;
    if (t < ni) {
        i = t + 1; j = 1;
    } else if (t < 2 * ni) {
        i = t - ni + 1; j = nj;
    } else if (t < 2 * ni + (nj - 2)) {
        i = 1; j = t - 2 * ni + 2;
    } else if (t < 2 * ni + 2 * (nj - 2)) {
        i = ni; j = t - 2 * ni - (nj - 2) + 2;
    } else {
        return;
    }
    si = i == 1 ? 2 : (i == ni ? ni - 1 : i);
    sj = j == 1 ? 2 : (j == nj ? nj - 1 : j);
    const size_t d = (size_t)(i - 1) + (size_t)(j - 1) * ni, s = (size_t)(si - 1) + (size_t)(sj - 1) * ni;
    a[d] = a[s];
    b[d] = b[s];
}

// ---- sweeps -------------------------------------------------------------------
// A sweep along x (X=true) walks i in row b=j; along y it walks j in column
// b=i.  Index a is the position along the sweep, na its extent.
template <bool X>
struct Line {
    const ClsGeo& G;
    int b;
    __device__ int na() const { return X ? G.ni : G.nj; }
    __device__ size_t c(int a) const { return X ? cix(G, a, b) : cix(G, b, a); }
    // masked face velocity ue(a,j) / ve(i,a) (Coupled_LS_VOF :778-797)
    // (both loads are issued before the flag selects: no dependent round trip)
    __device__ double vel(int a) const {
        const int af = a > 0 ? a : 1;
        if (X) {
            const double w = G.u[gix(G, a, b)];
            const bool on = G.uflag[cix(G, af, b)];
            return a == 0 || on ? w : 0.0;
        }
        const double w = G.v[gix(G, b, a)];
        const bool on = G.vflag[cix(G, b, af)];
        return a == 0 || on ? w : 0.0;
    }
    __device__ double ap(int a) const { return X ? G.aE[gix(G, a, b)] : G.aN[gix(G, b, a)]; }
    __device__ double am(int a) const { return X ? G.aW[gix(G, a, b)] : G.aS[gix(G, b, a)]; }
    __device__ double h(int a) const { return X ? G.dx[a - 1] : G.dy[a - 1]; }
    __device__ double hc() const { return X ? G.dy[b - 1] : G.dx[b - 1]; }
    __device__ double pos(int a) const { return X ? G.xc[a - 1] : G.yc[a - 1]; }
};

// rightflux/leftflux/topflux/bottomflux (:1416-1522): liquid volume swept
// through the far (side=+1) or near (side=-1) face over a width w
template <bool X>
__device__ double plic_face_flux(int side, double nxx, double nyy, double diss, double volf, double nxs, double nys,
                                 double phis, double vols, double hx, double hy, double w) {
    if (w == 0.0) return 0.0;
    const double nl = X ? nxx : nyy, ns = X ? nxs : nys, len = X ? hx : hy;
    const double sh = 0.5 * (w - len);
    const double dl = side > 0 ? diss - sh * nl : diss + sh * nl;
    double flux;
    if (vols < EPSI) {
        flux = X ? plic_frac(nxx, nyy, dl, w, hy) : plic_frac(nxx, nyy, dl, hx, w);
    } else if (vols >= 1.0 - EPSI) {
        return volf;
    } else {
        const double ps = side > 0 ? phis - sh * ns : phis + sh * ns;
        double volss;
        if (X)
            cell_geo(nxs, nys, ps, nxx, nyy, dl, w, hy, volss, flux);
        else
            cell_geo(nxs, nys, ps, nxx, nyy, dl, hx, w, volss, flux);
        if (volf / (1.0 - vols) >= 1.0 - VOFEPS) flux = 1.0 - volss;
        flux = flux / (1.0 - volss + TOLDENO);
        if (volf + vols > 1.0 - 1.0e-14) flux = 1.0;
    }
    return flux * w;
}

struct SweepIn {
    const double* vof;
    const double* phi;
    Planes pl;
    // PLIC fluxes precomputed by k_flux_pre: the sweep itself holds none of
    // the plane geometry, which keeps it free of scratch and light on registers
    const double* fr;
    const double* fl;
};

// the PLIC flux of vof_flux through face a with upwind cell a (far face of
// cell a) and through face a with upwind cell a+1 (near face of cell a+1)
template <bool X>
__device__ double plic_far(const Line<X>& L, const SweepIn& s, int a, size_t c, double f, double w, double dtv) {
    const ClsGeo& G = L.G;
    const double hx = X ? G.dx[a - 1] : G.dx[L.b - 1], hy = X ? G.dy[L.b - 1] : G.dy[a - 1];
    double nx, ny, d;
    plane_of(G, s.pl, c, f, hy, nx, ny, d);
    return plic_face_flux<X>(1, nx, ny, d, f, G.nxS[c], G.nyS[c], G.phiS[c], G.vofS[c], hx, hy, w * dtv);
}
template <bool X>
__device__ double plic_near(const Line<X>& L, const SweepIn& s, int a, size_t c, double f, double w, double dtv) {
    const ClsGeo& G = L.G;
    const double hx = X ? G.dx[a] : G.dx[L.b - 1], hy = X ? G.dy[L.b - 1] : G.dy[a];
    double nx, ny, d;
    plane_of(G, s.pl, c, f, hy, nx, ny, d);
    return -plic_face_flux<X>(-1, nx, ny, d, f, G.nxS[c], G.nyS[c], G.phiS[c], G.vofS[c], hx, hy, -w * dtv);
}

// vof_flux (X_Sweep :1014-1051) once the upwind cell of the face is known:
// f, vsU its vof and vofS, apU its aperture of the face, plic the PLIC flux
// precomputed for that side (read only for an interface cell, where
// VOFEPS < f < 1-VOFEPS)
__device__ __forceinline__ double upwind_flux(double w, double f, double vsU, double apU, const double* plic,
                                              size_t cU, double dtv) {
    double flux;
    if (f >= (1.0 - VOFEPS) || f <= VOFEPS) {
        flux = f * w * dtv;
        if (f >= 1.0) flux = w * dtv;
        if (f < 0.0) flux = 0.0;
    } else {
        flux = plic[cU];
    }
    if (apU > 1.0 - VOFEPS && vsU > 1.0 - VOFEPS) flux = 1.0 * w * dtv;
    return flux;
}

// level-set flux through interior face a (:1073-1096) from the level set
// at a-1, a, a+1, a+2 (pa1 = phi(a+1), pa2 = phi(a+2), ...)
__device__ __forceinline__ double ls_face(double w, int a, int n, double pm1, double p0, double p1, double p2,
                                          double dtv) {
    double l;
    if (w >= 0.0)
        l = (p1 + pm1) / 2.0;
    else if (a <= n - 2)
        l = (p2 + p0) / 2.0;
    else
        l = (p1 + pm1) / 2.0;
    return w * l * dtv;
}

// level-set flux through face 1 (:1098-1108 / :1218-1227)
template <bool X>
__device__ double ls_flux_first(const Line<X>& L, const double* phi, double dtv) {
    const double w = L.vel(1);
    const double p1 = phi[L.c(1)], p2 = phi[L.c(2)], p3 = phi[L.c(3)];
    double l;
    if (w >= 0.0)
        l = p1 + L.h(1) / 2.0 * (1.0 - w * dtv / L.h(1)) * (p2 - p1) / (L.pos(2) - L.pos(1));
    else
        l = p2 - L.h(2) / 2.0 * (1.0 + w * dtv / L.h(2)) * (p3 - p1) / (X ? 2.0 * L.h(2) : L.pos(3) - L.pos(1));
    return l * w * dtv;
}

// level-set flux through the last face (:1111-1123 / :1231-1243)
template <bool X>
__device__ double ls_flux_last(const Line<X>& L, const double* phi, double dtv) {
    const int n = L.na();
    const double w = L.vel(n);
    const double pn = phi[L.c(n)], pm = phi[L.c(n - 1)];
    double l;
    if (w > 0.0)
        l = pn + L.h(n) / 2.0 * (1.0 - w * dtv / L.h(n)) * (pn - pm) / (L.pos(n) - L.pos(n - 1));
    else
        l = pn - L.h(n - 1) / 2.0 * (1.0 + w * dtv / L.h(n - 1)) * (pn - pm) / (L.pos(n) - L.pos(n - 1));
    return l * w * dtv;
}

// The PLIC face fluxes of a sweep for the listed cells (VOFEPS < vof <
// 1-VOFEPS), packed one per lane: the far-face flux where the cell is upwind
// of its far face, the near-face flux where it is upwind of its near face --
// exactly the values vof_flux would compute for them in the sweep.
template <bool X>
__global__ void __launch_bounds__(256) k_flux_pre(ClsGeo G, SweepIn s, double dtv, const int* __restrict__ list,
                                                  const unsigned* __restrict__ cnt, double* __restrict__ fr,
                                                  double* __restrict__ fl) {
    const unsigned n = cnt[2];
    for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const int c = list[t];
        const double f = s.vof[c];
        if (f >= (1.0 - VOFEPS) || f <= VOFEPS) continue;  // not vof_flux's PLIC branch
        const int i = c % G.ni + 1, j = c / G.ni + 1;
        const Line<X> L{G, X ? j : i};
        const int a = X ? i : j;
        const double w = L.vel(a);
        if (X ? (w >= 0.0) : (w > 0.0)) fr[c] = plic_far<X>(L, s, a, c, f, w, dtv);
        const double w0 = L.vel(a - 1);
        if (!(X ? (w0 >= 0.0) : (w0 > 0.0))) fl[c] = plic_near<X>(L, s, a - 1, c, f, w0, dtv);
    }
}

// One sweep.  IMPL: first sweep of the sub-step (implicit divergence, clip);
// otherwise second sweep (explicit divergence with the first sweep's
// tv/tl).  level: Depthw+etau of the west inflow (x sweeps).  Every operand
// a cell may need is loaded up front, with neighbour indices clamped to the
// line, before any of them decides which one is used: the upwind choice
// then costs no second round trip to memory.
template <bool X, bool IMPL>
__global__ void __launch_bounds__(256) k_sweep(ClsGeo G, SweepIn s, double dtv, double level,
                                               const double* __restrict__ tv_in, const double* __restrict__ tl_in,
                                               double* __restrict__ tv_out, double* __restrict__ tl_out,
                                               double* __restrict__ vof_out, double* __restrict__ phi_out) {
    // x sweeps: XCD-aware tiles, so the tiles either side of a tile (whose
    // cells its stencil reads) share its L2; y sweeps: launch order, which
    // keeps vertically adjacent tiles on one XCD already
    int tx = blockIdx.x, ty = blockIdx.y;
    if (X) xcd_tile(tx, ty);
    const int i = tx * blockDim.x + threadIdx.x + 1;
    const int j = ty * blockDim.y + threadIdx.y + 1;
    if (i > G.ni || j > G.nj) return;
    const Line<X> L{G, X ? j : i};
    const int a = X ? i : j;
    const int n = L.na();
    const int am1 = a > 1 ? a - 1 : a, am2 = a > 2 ? a - 2 : am1;
    const int ap1 = a < n ? a + 1 : a, ap2 = a + 1 < n ? a + 2 : ap1;
    const size_t c = L.c(a), cm = L.c(am1), cp = L.c(ap1);
    const double vs = G.vofS[c], vsm = G.vofS[cm], vsp = G.vofS[cp];
    const double f0 = s.vof[c], fm = s.vof[cm], fp = s.vof[cp];
    const double wm = L.vel(a - 1), w0 = L.vel(a);
    const double am0 = L.am(a), ap0 = L.ap(a), apm = L.ap(am1), amp = L.am(ap1);
    const double ph0 = s.phi[c], phm1 = s.phi[cm], php1 = s.phi[cp];
    const double phm2 = s.phi[L.c(am2)], php2 = s.phi[L.c(ap2)];
    const double h = L.h(a);
    double tvi = 0.0, tli = 0.0;
    if (!IMPL) {
        tvi = tv_in[c];
        tli = tl_in[c];
    }
    // volume of fluid: temvf(a) after the sweep's in-place updates
    double t = f0;
    if (vs < 1.0 - VOFEPS) t = t / (1.0 - vs);
    if (a >= 2) {
        // face a-1: upwind cell a-1 or a
        const double fx = (X ? (wm >= 0.0) : (wm > 0.0)) ? upwind_flux(wm, fm, vsm, apm, s.fr, cm, dtv)
                                                         : upwind_flux(wm, f0, vs, am0, s.fl, c, dtv);
        t = t * (1.0 - vs) + fx * am0 / h;
    }
    {
        // face a: upwind cell a or a+1; the last face takes cell a's vof
        double fx;
        if (X ? (w0 >= 0.0) : (w0 > 0.0))
            fx = upwind_flux(w0, f0, vs, ap0, s.fr, c, dtv);
        else if (a < n)
            fx = upwind_flux(w0, fp, vsp, amp, s.fl, cp, dtv);
        else
            fx = f0 * w0 * dtv;
        t = t - fx * ap0 / h;
    }
    if (a == 1) {
        double fin = 0.0;
        if (X) {
            if (G.yc[j - 1] - G.dy[j - 1] / 2.0 < level) fin = f0 * wm * dtv;
        } else {
            fin = f0 * wm * dtv;
        }
        t = t + fin * L.am(1) / L.h(1);
    }
    // level set: temls(a)
    double q = ph0;
    if (a >= 3) q = q * (1.0 - vs) + ls_face(wm, a - 1, n, phm2, phm1, ph0, php1, dtv) * am0 / h;
    if (a >= 2 && a <= n - 1) q = q - ls_face(w0, a, n, phm1, ph0, php1, php2, dtv) * ap0 / h;
    if (a == 1) {
        q = q - ls_flux_first<X>(L, s.phi, dtv) * L.ap(1) / L.h(1);
        q = q + ph0 * wm * dtv * L.am(1) / L.h(1);
    }
    if (a == 2) q = q + ls_flux_first<X>(L, s.phi, dtv) * L.ap(1) / L.h(2);
    if (a == n) {
        const double fl = ls_flux_last<X>(L, s.phi, dtv);
        q = X ? q - fl * L.ap(n) / L.h(n) : q - fl / L.h(n);
    }
    const double dvel = w0 * ap0 - wm * am0;
    if (IMPL) {
        const double den = (1.0 - vs + TOLDENO) - dtv / h * dvel;
        t = t / den;
        q = q / den;
        tv_out[c] = t;
        tl_out[c] = q;
        double f = t * (1.0 - vs);
        if (f >= 1.0 - VOFEPS - vs) f = 1.0 - vs;
        if (f < VOFEPS) f = 0.0;
        vof_out[c] = f;
        phi_out[c] = q;
    } else {
        t = t + dtv / h * tvi * dvel;
        q = q + dtv / h * tli * dvel;
        vof_out[c] = t;
        phi_out[c] = vs < 1.0 - EPSI ? q / (1.0 - vs) : ph0;
    }
}

// ---- redistance -----------------------------------------------------------------

// isinsiderect (:1921-1955)
__device__ bool inside_rect(double xp, double yp, double dx, double dy) {
    const double x1 = -0.5 * dx, y1 = 0.5 * dy, x2 = 0.5 * dx, y2 = 0.5 * dy;
    const double x3 = 0.5 * dx, y3 = -0.5 * dy, x4 = -0.5 * dx, y4 = -0.5 * dy;
    const double b1 = sqrt((x1 - xp) * (x1 - xp) + (y1 - yp) * (y1 - yp));
    const double b2 = sqrt((x2 - xp) * (x2 - xp) + (y2 - yp) * (y2 - yp));
    const double b3 = sqrt((x3 - xp) * (x3 - xp) + (y3 - yp) * (y3 - yp));
    const double b4 = sqrt((x4 - xp) * (x4 - xp) + (y4 - yp) * (y4 - yp));
    const double u1 = 0.5 * (dx + b1 + b2), u2 = 0.5 * (dy + b2 + b3);
    const double u3 = 0.5 * (dx + b3 + b4), u4 = 0.5 * (dy + b4 + b1);
    const double a1 = sqrt(u1 * (u1 - dx) * (u1 - b1) * (u1 - b2));
    const double a2 = sqrt(u2 * (u2 - dy) * (u2 - b2) * (u2 - b3));
    const double a3 = sqrt(u3 * (u3 - dx) * (u3 - b3) * (u3 - b4));
    const double a4 = sqrt(u4 * (u4 - dy) * (u4 - b4) * (u4 - b1));
    return fabs(a1 + a2 + a3 + a4 - dx * dy) < 1.0e-10;
}

struct Target {
    double v, s, phi, x, y;  // vof, vofS, phi, centre of the target cell
};

// What one visit of a target by an interface source does (:1550-1686):
//   R2  two-phase target:   pa = sign(0.5-vof) * min(d, |pa|)
//   R3  three-phase target: if d < |pa|: pa = -d (liquid-full) or +d (empty)
//   RS  full-solid target:  if d < |pa|: pa = sign(dij) * d
// R2 and R3 reduce to a minimum over the visits in any order; RS depends on
// which source comes first among equal distances.
enum { V_NONE = 0, V_R2, V_R3, V_RS };
struct Visit {
    int kind;
    double d;
    double sg;  // RS sign
};

// the source cell's data a visit reads (its size, centre and solid line)
struct SrcCell {
    double dxs, dys, xc, yc, sxn, syn, sph, vs;
};
__device__ __forceinline__ SrcCell src_cell(const ClsGeo& G, int i, int j) {
    const size_t c = cix(G, i, j);
    return SrcCell{G.dx[i - 1], G.dy[j - 1], G.xc[i - 1], G.yc[j - 1], G.nxS[c], G.nyS[c], G.phiS[c], G.vofS[c]};
}

__device__ Visit propose(const SrcCell& S, int ii, int jj, double n1, double n2, double d1, const Target& T) {
    const double dxs = S.dxs, dys = S.dys;
    const double sxn = S.sxn, syn = S.syn, sph = S.sph;
    const int l = ii < -1 ? -1 : (ii > 1 ? 1 : ii);
    const int m = jj < -1 ? -1 : (jj > 1 ? 1 : jj);
    // point of the source cell's edge facing the target, moved off the solid
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
    const double ddx = T.x - S.xc, ddy = T.y - S.yc;
    const int rule = T.s < EPSI ? V_R2 : V_R3;
    const double dv = n1 * xv + n2 * yv + d1;
    if (dv * fsign(1.0, T.phi) <= 0.0) {
        const double ax = ddx - xv, ay = ddy - yv;
        return {rule, sqrt(ax * ax + ay * ay), 0.0};
    }
    // projection of the target centre onto the source's PLIC line
    const double dij = n1 * ddx + n2 * ddy + d1;
    const double xp = ddx - dij * n1, yp = ddy - dij * n2;
    const bool inside = inside_rect(xp, yp, dxs, dys);
    if (inside && T.s < 1.0 - EPSI) {
        const double ax = ddx - xp, ay = ddy - yp;
        return {rule, sqrt(ax * ax + ay * ay), 0.0};
    }
    if (T.s >= 1.0 - EPSI) {
        if (S.vs < EPSI) {
            const double ax = ddx - xp, ay = ddy - yp;
            return {V_RS, sqrt(ax * ax + ay * ay), fsign(1.0, dij)};
        }
        return {V_NONE, 0.0, 0.0};
    }
    // nearest end of the segment
    const double tol = 1.0e-24;
    const double xoff = fmax(fabs(xp) - 0.5 * dxs, 0.0);
    const double yoff = fmax(fabs(yp) - 0.5 * dys, 0.0);
    const double xfc = fsign(1.0, xp) * 0.5 * dxs;
    const double yfc = fsign(1.0, yp) * 0.5 * dys;
    double xs, ys;
    if (xoff * fabs(n1 + tol) >= yoff * fabs(n2 + tol)) {
        xs = xfc;
        ys = (d1 + n1 * xs) / (-n2 + tol);
    } else {
        ys = yfc;
        xs = (d1 + n2 * ys) / (-n1 + tol);
    }
    const double ax = ddx - xs, ay = ddy - ys;
    const double deuc = sqrt(ax * ax + ay * ay);
    if (sxn * xs + syn * ys + sph >= 0.0) return {rule, deuc, 0.0};
    return {V_NONE, 0.0, 0.0};
}

__device__ __forceinline__ void apply(const Visit& w, const Target& T, double& pa) {
    if (w.kind == V_R2) {
        pa = fsign(1.0, 0.5 - T.v) * fmin(w.d, fabs(pa));
    } else if (w.kind == V_R3) {
        if (fabs(w.d) < fabs(pa) && T.v >= 1.0 - VOFEPS - T.s) pa = -fabs(w.d);
        if (fabs(w.d) < fabs(pa) && T.v < VOFEPS) pa = fabs(w.d);
    } else if (w.kind == V_RS) {
        if (fabs(w.d) < fabs(pa)) pa = w.sg * w.d;
    }
}

constexpr unsigned long long BEST_NONE = 0x7FF0000000000000ull;  // +inf: no proposal yet

__device__ __forceinline__ bool rd_source(double v, double s) { return (0.0 + VOFEPS) < v && v < (1.0 - VOFEPS - s); }
__device__ __forceinline__ bool rd_target(double v, double s) { return v < 0.0 + VOFEPS || v >= 1.0 - VOFEPS - s; }

// (source, offset) pairs of the listed interface cells; pair p is source
// p/81 and window offset o = p%81: column offset o%9 - BAND, row offset
// o/9 - BAND (consecutive lanes walk along a row of the window)
struct PairView {
    const int* srcs;
    const unsigned* cnt;
};

// Each pair proposes a distance to its target; two- and three-phase
// targets keep the minimum (the bit pattern of a non-negative double orders
// like the value).  Full-solid targets are left to the ordered gather of
// k_redistance_band.  `best` and `hit` hold +inf and 0 on entry: every cell
// a pair can reach lies in the band, and k_redistance_band puts them back
// after reading them.
__global__ void __launch_bounds__(256) k_rd_scatter(ClsGeo G, const double* __restrict__ vof,
                                                    const double* __restrict__ phi, Planes P, PairView pv,
                                                    unsigned long long* __restrict__ best,
                                                    unsigned char* __restrict__ hit) {
    const unsigned long long np = (unsigned long long)pv.cnt[0] * 81ull;
    for (unsigned long long p = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; p < np;
         p += (unsigned long long)gridDim.x * blockDim.x) {
        const int c0 = pv.srcs[p / 81];
        const int o = (int)(p % 81);
        const int si = c0 % G.ni + 1, sj = c0 / G.ni + 1;
        const int ii = o % 9 - BAND, jj = o / 9 - BAND;
        const int ti = si + ii, tj = sj + jj;
        if (ti < 1 || ti > G.ni || tj < 1 || tj > G.nj) continue;
        const size_t t = cix(G, ti, tj);
        // the target's and the source's data, loaded before the tests that
        // decide whether they are used
        Target T;
        T.v = vof[t];
        T.s = G.vofS[t];
        T.phi = phi[t];
        T.x = G.xc[ti - 1];
        T.y = G.yc[tj - 1];
        const SrcCell S = src_cell(G, si, sj);
        const double sv = vof[c0], snx = P.nx[c0], sny = P.ny[c0], sdis = P.dis[c0];
        if (!rd_target(T.v, T.s) || T.s >= 1.0 - EPSI) continue;
        double n1, n2, d1;
        plane_pick(sv, S.vs, S.dys, snx, sny, sdis, n1, n2, d1);
        const Visit w = propose(S, ii, jj, n1, n2, d1, T);
        if (w.kind == V_NONE) continue;
        if (w.kind == V_R2) hit[t] = 1;
        if (!(w.d != w.d)) atomicMin(best + t, (unsigned long long)__double_as_longlong(w.d));
    }
}

// The band bitmap of k_redistance: word (w, j) is nonzero when an interface
// cell lies in rows j-BAND..j+BAND of mask words w-1..w+1, a superset of the
// windows of the 64 cells of word (w, j) (BAND < 64)
__global__ void __launch_bounds__(256) k_rd_dilate(int nj, int mw, const unsigned long long* __restrict__ mask,
                                                   unsigned long long* __restrict__ dil, int* __restrict__ band,
                                                   unsigned* __restrict__ nband) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= mw * nj) return;
    const int w = k % mw, j = k / mw;
    unsigned long long any = 0ull;
    for (int r = max(j - BAND, 0); r <= min(j + BAND, nj - 1); ++r) {
        const unsigned long long* row = mask + (size_t)r * mw;
        any |= row[w];
        if (w > 0) any |= row[w - 1];
        if (w + 1 < mw) any |= row[w + 1];
    }
    dil[k] = any;
    if (any) band[atomicAdd(nband, 1u)] = k;
}

// Redistance, per cell: interface cells take their plane constant, cells
// reached from an interface cell the nearest distance, the rest the band
// value.  Full-solid targets replay their visits in the reference's source
// order (i ascending, then j ascending).  `near`: the cell's band bitmap
// word is nonzero (else no interface cell lies in its window).
__device__ void redistance_cell(const ClsGeo& G, const double* __restrict__ vof, const Planes& P,
                                const unsigned long long* __restrict__ mask, int mw,
                                const unsigned long long* __restrict__ best, const unsigned char* __restrict__ hit,
                                double* __restrict__ phi, int ti, int tj, bool near) {
    const size_t c = cix(G, ti, tj);
    Target T;
    T.v = vof[c];
    T.s = G.vofS[c];
    // the scatter's minimum and hit flag, loaded with the cell's own data
    // rather than after the window scan that decides whether they are used
    const unsigned long long bc = best[c];
    const bool hc = hit[c] != 0;
    const double hy = G.dy[tj - 1];
    if (rd_source(T.v, T.s)) {
        double n1, n2, d1;
        plane_of(G, P, c, T.v, hy, n1, n2, d1);
        phi[c] = d1;  // an interface cell keeps its own plane constant
        return;
    }
    double pa = 1.0e4;
    bool visited = false;
    if (rd_target(T.v, T.s) && near) {
        const int ilo = max(ti - BAND, 1), ihi = min(ti + BAND, G.ni);
        const int jlo = max(tj - BAND, 1), jhi = min(tj + BAND, G.nj);
        // window bits per row: bit k <-> source column ilo+k
        unsigned rows[2 * BAND + 1];
        unsigned any = 0;
        for (int sj = jlo; sj <= jhi; ++sj) {
            const unsigned long long* r = mask + (size_t)(sj - 1) * mw;
            const int b0 = ilo - 1, w0 = b0 >> 6, w1 = (ihi - 1) >> 6, sh = b0 & 63;
            unsigned long long bits = r[w0] >> sh;
            if (w1 != w0 && sh) bits |= r[w1] << (64 - sh);
            const unsigned win = (unsigned)(bits & ((1ull << (ihi - ilo + 1)) - 1ull));
            rows[sj - jlo] = win;
            any |= win;
        }
        visited = any != 0;
        if (visited) {
            if (T.s >= 1.0 - EPSI) {
                T.phi = phi[c];
                T.x = G.xc[ti - 1];
                T.y = G.yc[tj - 1];
                for (int si = ilo; si <= ihi; ++si) {
                    for (int sj = jlo; sj <= jhi; ++sj) {
                        if (!((rows[sj - jlo] >> (si - ilo)) & 1u)) continue;
                        const size_t sc = cix(G, si, sj);
                        double n1, n2, d1;
                        plane_of(G, P, sc, vof[sc], G.dy[sj - 1], n1, n2, d1);
                        apply(propose(src_cell(G, si, sj), ti - si, tj - sj, n1, n2, d1, T), T, pa);
                    }
                }
            } else {
                const double b = __longlong_as_double((long long)bc);
                if (T.s < EPSI) {
                    if (hc) pa = fsign(1.0, 0.5 - T.v) * fmin(b, 1.0e4);
                } else if (b < 1.0e4) {
                    pa = T.v >= 1.0 - VOFEPS - T.s ? -b : b;
                }
            }
        }
    }
    if (visited && !(pa > 9.999e3)) {
        phi[c] = pa;
        return;
    }
    const double hx = G.dx[ti - 1];
    if (T.v <= VOFEPS)
        phi[c] = (double)(BAND + 1) * sqrt(hx * hx + hy * hy);
    else if (T.v >= (1.0 - VOFEPS - T.s))
        phi[c] = -(double)(BAND + 1) * sqrt(hx * hx + hy * hy);
}

// the cells of the listed band words, one wavefront per word (the window
// scans and nearest-distance gathers; few words, many registers).  Each
// cell's minimum and hit flag go back to +inf and 0 once read, ready for
// the next k_rd_scatter: the band holds every cell a pair can reach.
__global__ void __launch_bounds__(256) k_redistance_band(ClsGeo G, const double* __restrict__ vof, Planes P,
                                                         const unsigned long long* __restrict__ mask, int mw,
                                                         const int* __restrict__ band,
                                                         const unsigned* __restrict__ nband,
                                                         unsigned long long* __restrict__ best,
                                                         unsigned char* __restrict__ hit,
                                                         double* __restrict__ phi) {
    const unsigned n = *nband;
    const int lane = threadIdx.x & 63;
    for (unsigned e = blockIdx.x * 4 + (threadIdx.x >> 6); e < n; e += gridDim.x * 4) {
        const int k = band[e];
        const int ti = (k % mw) * 64 + lane + 1, tj = k / mw + 1;
        if (ti <= G.ni) {
            redistance_cell(G, vof, P, mask, mw, best, hit, phi, ti, tj, true);
            const size_t c = cix(G, ti, tj);
            best[c] = BEST_NONE;
            hit[c] = 0;
        }
    }
}

// every other cell: no interface cell in its window, so it takes the band
// value (or keeps its own level set), with the occupancy of a light kernel
__global__ void __launch_bounds__(256) k_redistance_far(ClsGeo G, const double* __restrict__ vof,
                                                        const unsigned long long* __restrict__ dil, int mw,
                                                        double* __restrict__ phi) {
    const int ti = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int tj = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (ti > G.ni || tj > G.nj) return;
    if (dil[(size_t)(tj - 1) * mw + ((ti - 1) >> 6)] != 0ull) return;
    const size_t c = cix(G, ti, tj);
    const double v = vof[c], s = G.vofS[c];
    const double hx = G.dx[ti - 1], hy = G.dy[tj - 1];
    if (v <= VOFEPS)
        phi[c] = (double)(BAND + 1) * sqrt(hx * hx + hy * hy);
    else if (v >= (1.0 - VOFEPS - s))
        phi[c] = -(double)(BAND + 1) * sqrt(hx * hx + hy * hy);
}

__global__ void k_fill_u64(unsigned long long* __restrict__ a, size_t n, unsigned long long v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

__global__ void k_zero(unsigned* __restrict__ a, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) a[t] = 0u;
}

// ---- interface on the U / V cells (Initial_ClsVofUV :402-710) ----------------

// LagrangePoly (:712-726) at PosTar = 0
__device__ double lagrange4(const double (&posi)[4], const double (&valu)[4]) {
    double val = 0.0;
    for (int i = 0; i < 4; ++i) {
        double mult = valu[i];
        for (int j = 0; j < 4; ++j)
            if (j != i) mult = mult * (0.0 - posi[j]) / (posi[i] - posi[j]);
        val = val + mult;
    }
    return val;
}

struct UVIn {
    const double *vof, *phi, *nx, *ny;   // P cells
    const double *vofST, *phiST;         // staggered solid fraction / level set
};

// liquid fraction of the half of P cell (i,j) facing the staggered cell
// (right: the half of the next cell); false when no branch of :448-536 /
// :591-672 assigns it
template <int W>
__device__ bool half_volf(const ClsGeo& G, const UVIn& in, bool right, int i, int j, double phisT, double& out) {
    const size_t c = cix(G, i, j);
    const double vf = in.vof[c], vs = G.vofS[c];
    const double dx = G.dx[i - 1], dy = G.dy[j - 1];
    double ph;
    if (W == 0)
        ph = right ? in.phi[c] - dx / 4.0 * in.nx[c] : in.phi[c] + dx / 4.0 * in.nx[c];
    else
        ph = right ? in.phi[c] - dy / 4.0 * in.ny[c] : in.phi[c] + dy / 4.0 * in.ny[c];
    const double hx = W == 0 ? dx / 2.0 : dx, hy = W == 0 ? dy : dy / 2.0;
    bool set = false;
    double v = 0.0;
    if (vf >= 1.0 - VOFEPS && vs < EPSI) { v = 1.0; set = true; }
    if (vf < 1.0 - VOFEPS && vf >= VOFEPS && vs < EPSI) { v = plic_frac(in.nx[c], in.ny[c], ph, hx, hy); set = true; }
    if (vf < VOFEPS && vs < EPSI) { v = 0.0; set = true; }
    if (vf < VOFEPS && vs >= 1.0 - EPSI) { v = 0.0; set = true; }
    if (vf < VOFEPS && vs < (W == 0 ? 1.0 - VOFEPS : 1.0 - EPSI) && vs >= EPSI) { v = 0.0; set = true; }
    if (vf >= VOFEPS && vf < 1.0 - VOFEPS && vs >= EPSI && vs < 1.0 - EPSI) {
        double vols;
        cell_geo(G.nxS[c], G.nyS[c], phisT, in.nx[c], in.ny[c], ph, hx, hy, vols, v);
        set = true;
    }
    if (set) out = v;
    return set;
}

// previous staggered cell in the reference's loop order (U: j outer, i
// inner over i < ni; V: i outer, j inner over j < nj); false before the first
template <int W>
__device__ __forceinline__ bool prev_cell(const ClsGeo& G, int& i, int& j) {
    if (W == 0) {
        if (i > 1) { --i; return true; }
        if (j > 1) { --j; i = G.ni - 1; return true; }
        return false;
    }
    if (j > 1) { --j; return true; }
    if (i > 1) { --i; j = G.nj - 1; return true; }
    return false;
}

// volfl / volfr of staggered cell (i,j), walking back to the last cell that
// assigned it when this one does not (stale locals of the reference)
template <int W>
__device__ double half_stale(const ClsGeo& G, const UVIn& in, bool right, int i, int j) {
    double v = 0.0;
    for (;;) {
        const int ip = W == 0 ? i + 1 : i, jp = W == 0 ? j : j + 1;
        const double phisT = in.phiST[cix(G, i, j)];
        if (right ? half_volf<W>(G, in, true, ip, jp, phisT, v) : half_volf<W>(G, in, false, i, j, phisT, v)) return v;
        if (!prev_cell<W>(G, i, j)) return 0.0;
    }
}

template <int W>
__device__ void uv_cell(const ClsGeo& G, const UVIn& in, int i, int j, double& tv, double& tp, double& tx, double& ty) {
    const int ip = W == 0 ? i + 1 : i, jp = W == 0 ? j : j + 1;
    const bool lagr = W == 0 ? (i >= 3 && i <= G.ni - 2) : (j >= 3 && j <= G.nj - 2);
    double p, a, b;
    if (lagr) {
        double posi[4], v1[4], v2[4], v3[4];
        for (int m = 0; m < 4; ++m) {
            const int ii = W == 0 ? i - 1 + m : i, jj = W == 0 ? j : j - 1 + m;
            posi[m] = W == 0 ? (G.xc[i - 1] + 0.5 * G.dx[i - 1]) - G.xc[ii - 1]
                             : (G.yc[j - 1] + 0.5 * G.dy[j - 1]) - G.yc[jj - 1];
            const size_t c = cix(G, ii, jj);
            v1[m] = in.phi[c];
            v2[m] = in.nx[c];
            v3[m] = in.ny[c];
        }
        p = lagrange4(posi, v1);
        a = lagrange4(posi, v2);
        b = lagrange4(posi, v3);
    } else {
        const size_t c0 = cix(G, i, j), c1 = cix(G, ip, jp);
        p = 0.5 * (in.phi[c1] + in.phi[c0]);
        a = 0.5 * (in.nx[c1] + in.nx[c0]);
        b = 0.5 * (in.ny[c1] + in.ny[c0]);
    }
    const double temn = W == 0 ? sqrt(a * a + b * b) + 1.0e-14 : sqrt(a * a + b * b + 1.0e-30);
    tp = p;
    tx = a / temn;
    ty = b / temn;
    const double vl = half_stale<W>(G, in, false, i, j);
    const double vr = half_stale<W>(G, in, true, i, j);
    double t = 0.5 * (vl + vr);
    const double vsT = in.vofST[cix(G, i, j)];
    if (t < VOFEPS) t = 0.0;
    if (t > 1.0 - VOFEPS - vsT) t = 1.0 - vsT;
    const int im = W == 0 ? i - 1 : i, jm = W == 0 ? j : j - 1;
    if ((W == 0 ? i > 1 : j > 1) && in.vof[cix(G, im, jm)] >= 1.0 - VOFEPS - G.vofS[cix(G, im, jm)] &&
        in.vof[cix(G, ip, jp)] >= 1.0 - VOFEPS - G.vofS[cix(G, ip, jp)])
        t = 1.0 - vsT;
    tv = t;
}

// one thread per staggered cell; the last column (U) / row (V) copies its
// neighbour (:555-558 / :700-703)
template <int W>
__global__ void __launch_bounds__(256) k_to_uv(ClsGeo G, UVIn in, double* __restrict__ tvof,
                                               double* __restrict__ tphi, double* __restrict__ tnx,
                                               double* __restrict__ tny) {
    // U: XCD-aware tiles (the stencil reaches along x); V: launch order
    int bx = blockIdx.x, by = blockIdx.y;
    if (W == 0) xcd_tile(bx, by);
    const int i = bx * blockDim.x + threadIdx.x + 1;
    const int j = by * blockDim.y + threadIdx.y + 1;
    if (i > G.ni || j > G.nj) return;
    const bool last = W == 0 ? i == G.ni : j == G.nj;
    const int si = last && W == 0 ? i - 1 : i, sj = last && W == 1 ? j - 1 : j;
    double tv, tp, tx, ty;
    uv_cell<W>(G, in, si, sj, tv, tp, tx, ty);
    const size_t c = cix(G, i, j);
    tvof[c] = tv;
    tphi[c] = tp;
    tnx[c] = tx;
    tny[c] = ty;
}

ClsGeo make_geo(Ctx* c) {
    ClsGeo G;
    G.ni = c->ni;
    G.nj = c->nj;
    G.dx = c->dx;
    G.dy = c->dy;
    G.xc = c->cls.xc;
    G.yc = c->cls.yc;
    G.vofS = (const double*)c->field[SG_F_VOFS];
    G.nxS = (const double*)c->field[SG_F_NXS];
    G.nyS = (const double*)c->field[SG_F_NYS];
    G.phiS = (const double*)c->field[SG_F_PHIS];
    G.aE = (const double*)c->field[SG_F_AE];
    G.aW = (const double*)c->field[SG_F_AW];
    G.aN = (const double*)c->field[SG_F_AN];
    G.aS = (const double*)c->field[SG_F_AS];
    G.uflag = (const int*)c->field[SG_F_UFLAG];
    G.vflag = (const int*)c->field[SG_F_VFLAG];
    G.u = (const double*)c->field[SG_F_U];
    G.v = (const double*)c->field[SG_F_V];
    return G;
}

}  // namespace

int clsvof_allocate(Ctx* c) {
    const size_t n = (size_t)c->ni * c->nj;
    ClsBuffers& B = c->cls;
    B.mw = (c->ni + 63) / 64;
    B.xc = (double*)dev_alloc(c, c->ni * sizeof(double));
    B.yc = (double*)dev_alloc(c, c->nj * sizeof(double));
    B.dis = (double*)dev_alloc(c, n * sizeof(double));
    B.vt = (double*)dev_alloc(c, n * sizeof(double));
    B.pt = (double*)dev_alloc(c, n * sizeof(double));
    B.tv = (double*)dev_alloc(c, n * sizeof(double));
    B.tl = (double*)dev_alloc(c, n * sizeof(double));
    B.stale = (unsigned char*)dev_alloc(c, n);
    B.hit = (unsigned char*)dev_alloc(c, n);
    B.best = (unsigned long long*)dev_alloc(c, n * sizeof(unsigned long long));
    if (B.best) {
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, B.best, n,
                           BEST_NONE);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) return SG_ERR_ALLOC;
    }
    B.srcs = (int*)dev_alloc(c, n * sizeof(int));
    B.stales = (int*)dev_alloc(c, n * sizeof(int));
    B.ifc = (int*)dev_alloc(c, n * sizeof(int));
    B.fr = (double*)dev_alloc(c, n * sizeof(double));
    B.fl = (double*)dev_alloc(c, n * sizeof(double));
    B.cnt = (unsigned*)dev_alloc(c, ClsBuffers::CNT_SLOTS * sizeof(unsigned));
    B.mask = (unsigned long long*)dev_alloc(c, (size_t)B.mw * c->nj * sizeof(unsigned long long));
    B.dil = (unsigned long long*)dev_alloc(c, (size_t)B.mw * c->nj * sizeof(unsigned long long));
    B.band = (int*)dev_alloc(c, (size_t)B.mw * c->nj * sizeof(int));
    if (!B.xc || !B.yc || !B.dis || !B.vt || !B.pt || !B.tv || !B.tl || !B.stale || !B.hit || !B.best || !B.srcs ||
        !B.stales || !B.ifc || !B.fr || !B.fl || !B.cnt || !B.mask || !B.dil || !B.band)
        return SG_ERR_ALLOC;
    return SG_OK;
}

// the recorded sweep launches' durations into the totals (waits for them)
int sweep_timing_drain(Ctx* c) {
    auto& T = c->sweep_timing;
    for (int k = 0; k < T.used; ++k) {
        SG_HIP(hipEventSynchronize(T.ev[2 * k + 1]));
        float ms = 0.f;
        SG_HIP(hipEventElapsedTime(&ms, T.ev[2 * k], T.ev[2 * k + 1]));
        T.ms += ms;
        T.launches += 1;
    }
    T.used = 0;
    return SG_OK;
}

// Coupled_LS_VOF on the device fields VOF, PHI (in/out), NX, NY (out),
// U, V (in).  `levels` holds Depthw+etau of each sub-step.
int clsvof_advect(Ctx* c, double dt, int nsub, const double* levels) {
    const ClsGeo G = make_geo(c);
    ClsBuffers& B = c->cls;
    double* vof = (double*)c->field[SG_F_VOF];
    double* phi = (double*)c->field[SG_F_PHI];
    const Planes P{(double*)c->field[SG_F_NX], (double*)c->field[SG_F_NY], B.dis};
    const dim3 blk(64, 4), grd((c->ni + 63) / 64, (c->nj + 3) / 4);
    const int ring = 2 * c->ni + 2 * (c->nj - 2);
    const dim3 rblk(256), rgrd((ring + 255) / 256);
    const dim3 pblk(256), pgrd(1024);  // grid-stride loops over the listed pairs / cells
    const dim3 rcblk(64, RECON_TY),
        rcgrd((c->ni + 63) / 64, (c->nj + RECON_TY * RECON_R - 1) / (RECON_TY * RECON_R));
    // sweeps: 32x8 tiles (3% faster per transport call than 64x4, 256x1 or
    // 128x2 at 2048^2: the y stencil's rows stay within fewer tiles)
    const dim3 sblk(32, 8), sgrd((c->ni + 31) / 32, (c->nj + 7) / 8);
    const double dtv = dt / (double)nsub;
    hipStream_t st = c->stream;
    // three counters per reconstruction, three reconstructions per sub-step,
    // then one band-word count per sub-step
    const int slots = 10 * nsub;
    if (slots > ClsBuffers::CNT_SLOTS) {
        set_error("clsvof_advect: too many sub-steps");
        return SG_ERR_ARG;
    }
    hipLaunchKernelGGL(k_zero, dim3((slots + 255) / 256), dim3(256), 0, st, B.cnt, slots);
    int slot = 0;
    // `scan`: the interface list of the last full-grid scan, when vof and
    // vofS have not changed since (re-distancing changes phi only)
    // last: the sub-step's final reconstruction (re-distancing sources);
    // full: also n = (0,1) on the other cells -- only the transport's very
    // last one, since within the transport a stored plane is read only
    // where the cell is an interface cell of the current vof (plane_of,
    // plane_pick), and every such plane is rewritten by k_recon_ifc
    auto recon = [&](const double* v, const double* p, bool last, const unsigned* scan, bool full) {
        unsigned* cnt = B.cnt + slot;
        slot += 3;
        if (!scan)
            hipLaunchKernelGGL(k_recon, rcgrd, rcblk, 0, st, G, v, P, full ? 1 : 0,
                               last ? B.mask : (unsigned long long*)nullptr, B.mw, B.srcs, B.ifc, cnt);
        const unsigned* nifc = scan ? scan + 2 : cnt + 2;
        hipLaunchKernelGGL(k_recon_ifc, dim3(256), pblk, 0, st, G, v, p, P, B.stale, B.ifc, nifc, B.stales, cnt);
        hipLaunchKernelGGL(k_dis_fix, dim3(64), pblk, 0, st, G, v, B.stale, B.stales, cnt, B.dis);
        // the interface count (entry 2) the flux precomputation reads
        return scan ? scan : (const unsigned*)cnt;
    };
    const unsigned* last_scan = nullptr;  // the previous sub-step's final scan
    // sg_sweep_timing: an event before (end = 0) and after (end = 1) each sweep
    auto& TM = c->sweep_timing;
    if (TM.on && TM.used + 2 * nsub > Ctx::SweepTiming::POOL) {
        const int rc = sweep_timing_drain(c);
        if (rc) return rc;
    }
    auto sweep_mark = [&](int end) -> hipError_t {
        if (!TM.on) return hipSuccess;
        const hipError_t e = hipEventRecord(TM.ev[2 * TM.used + end], st);
        if (end) ++TM.used;
        return e;
    };
    for (int k = 1; k <= nsub; ++k) {
        const double level = levels[k - 1];
        // first sweep: (vof, phi) -> (vt, pt), implicit divergence; vof is
        // the one the previous sub-step's last reconstruction scanned
        const unsigned* c1 = recon(vof, phi, false, last_scan, false);
        const SweepIn s1{vof, phi, P, B.fr, B.fl};
        if (k % 2 == 1)
            hipLaunchKernelGGL((k_flux_pre<true>), pgrd, pblk, 0, st, G, s1, dtv, B.ifc, c1, B.fr, B.fl);
        else
            hipLaunchKernelGGL((k_flux_pre<false>), pgrd, pblk, 0, st, G, s1, dtv, B.ifc, c1, B.fr, B.fl);
        SG_HIP(sweep_mark(0));
        if (k % 2 == 1)
            hipLaunchKernelGGL((k_sweep<true, true>), sgrd, sblk, 0, st, G, s1, dtv, level, nullptr, nullptr, B.tv,
                               B.tl, B.vt, B.pt);
        else
            hipLaunchKernelGGL((k_sweep<false, true>), sgrd, sblk, 0, st, G, s1, dtv, level, nullptr, nullptr, B.tv,
                               B.tl, B.vt, B.pt);
        SG_HIP(sweep_mark(1));
        hipLaunchKernelGGL(k_bc, rgrd, rblk, 0, st, c->ni, c->nj, B.vt, B.pt);
        // second sweep: (vt, pt) -> (vof, phi), explicit divergence
        const unsigned* c2 = recon(B.vt, B.pt, false, nullptr, false);
        const SweepIn s2{B.vt, B.pt, P, B.fr, B.fl};
        if (k % 2 == 1)
            hipLaunchKernelGGL((k_flux_pre<false>), pgrd, pblk, 0, st, G, s2, dtv, B.ifc, c2, B.fr, B.fl);
        else
            hipLaunchKernelGGL((k_flux_pre<true>), pgrd, pblk, 0, st, G, s2, dtv, B.ifc, c2, B.fr, B.fl);
        SG_HIP(sweep_mark(0));
        if (k % 2 == 1)
            hipLaunchKernelGGL((k_sweep<false, false>), sgrd, sblk, 0, st, G, s2, dtv, level, B.tv, B.tl, nullptr,
                               nullptr, vof, phi);
        else
            hipLaunchKernelGGL((k_sweep<true, false>), sgrd, sblk, 0, st, G, s2, dtv, level, B.tv, B.tl, nullptr,
                               nullptr, vof, phi);
        SG_HIP(sweep_mark(1));
        if (c->moving.shape.kind != SG_SOLID_NONE) {
            // ObjectMovement (Clsvof_Mod.f90:2384-2398) over the sub-step,
            // then SolidVolumeFraction on the P cells (:920-921)
            auto& m = c->moving;
            m.shape.cy = m.shape.cy + 0.5 * m.asy * (dtv * dtv) + m.vs * dtv;
            m.shape.cx = m.shape.cx + 0.5 * m.asx * (dtv * dtv) + m.us * dtv;
            m.us = m.us + m.asx * dtv;
            m.vs = m.vs + m.asy * dtv;
            const int rc = solid_fraction_step(c, &m.shape);
            if (rc) return rc;
        }
        hipLaunchKernelGGL(k_bc, rgrd, rblk, 0, st, c->ni, c->nj, vof, phi);
        // reconstruction (full normals out) and re-distancing
        const unsigned* cnt = recon(vof, phi, true, nullptr, k == nsub);
        last_scan = cnt;
        const PairView pv{B.srcs, cnt};
        hipLaunchKernelGGL(k_rd_scatter, pgrd, pblk, 0, st, G, vof, phi, P, pv, B.best, B.hit);
        unsigned* nband = B.cnt + 9 * nsub + (k - 1);
        hipLaunchKernelGGL(k_rd_dilate, dim3((unsigned)((B.mw * c->nj + 255) / 256)), dim3(256), 0, st, c->nj, B.mw,
                           (const unsigned long long*)B.mask, B.dil, B.band, nband);
        hipLaunchKernelGGL(k_redistance_far, grd, blk, 0, st, G, vof, (const unsigned long long*)B.dil, B.mw, phi);
        hipLaunchKernelGGL(k_redistance_band, dim3(2048), dim3(256), 0, st, G, vof, P, B.mask, B.mw,
                           (const int*)B.band, (const unsigned*)nband, B.best, B.hit, phi);
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace sg

namespace sg {

// Initial_ClsVofUV for both staggered grids: device fields VOF, PHI, NX, NY
// (+ the solid geometry) -> UVOF/UPHI/UNX/UNY and VVOF/VPHI/VNX/VNY.
int interface_to_uv(Ctx* c) {
    const ClsGeo G = make_geo(c);
    const dim3 blk(64, 4), grd((c->ni + 63) / 64, (c->nj + 3) / 4);
    auto f = [&](int k) { return (double*)c->field[k]; };
    UVIn in{f(SG_F_VOF), f(SG_F_PHI), f(SG_F_NX), f(SG_F_NY), c->stag[0].a[STG_VOFS], c->stag[0].a[STG_PHIS]};
    hipLaunchKernelGGL(k_to_uv<0>, grd, blk, 0, c->stream, G, in, f(SG_F_UVOF), f(SG_F_UPHI), f(SG_F_UNX),
                       f(SG_F_UNY));
    in.vofST = c->stag[1].a[STG_VOFS];
    in.phiST = c->stag[1].a[STG_PHIS];
    hipLaunchKernelGGL(k_to_uv<1>, grd, blk, 0, c->stream, G, in, f(SG_F_VVOF), f(SG_F_VPHI), f(SG_F_VNX),
                       f(SG_F_VNY));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace sg
