// Cut-cell geometry rebuilt on the device for a solid given as a shape: the
// solid level set and normal on the P, U and V cells and their solid
// fractions (SolidVolumeFraction's role, Clsvof_Mod.f90:2553-2929, for a
// box, circle, ellipse or the oil boom's cylinder with its bar,
// Main.f90:173-182), then Grid_Preprocess and NewCellFace as the reference
// calls them each step (Solver_Mod.f90:303-304): Cell_Geo_Cal /
// Edge_Geo_Cal (Cutcell_Mod.f90:34-337: edge fractions and face centres,
// harmonic face averaging, the closing loop, wall lengths, the corner patch
// of ICorProb), DefineMomentumExchangeCell (:360-397), NumberExternalCell
// (:340-358), NewCellFace (:399-549) and EastFaceInterpolationInf /
// NorthFaceInterpolationInf (:551-700).  Every kernel evaluates the
// reference's expressions in its order with FP contraction off.
//
// Steps of the reference that are sequential in loop order are recast:
//  * the closing loop visits cells i outer, j inner; a cell's decision
//    depends on the decisions of its west and south neighbours (the cells
//    visited before it that close one of its faces) and on its own averaged
//    east and north faces, so the decisions solve a recurrence over an
//    acyclic graph, found by repeating a parallel pass until nothing changes;
//  * the harmonic averaging, DefineMomentumExchangeCell, the corner patch
//    and NewCellFace read only values no other cell of the same loop writes,
//    so their cells are independent;
//  * AlE(0,j) = AlN(1,j) (:606) reads the north factor the previous
//    geometry left, so the east factors are formed before the north ones.
// Storage the reference reads without defining it: aperture ghosts 0;
// Posnu of the U column i = Isize and the V row j = Jsize (never numbered)
// a fresh allocation's 0, i.e. numbered; face-factor ghosts neither
// interpolation routine writes the nearest written entry.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr double G_VOFEPS = 1.0e-14;   // Clsvof_Mod.f90:10
constexpr double G_EPSIF = 1.0e-2;     // StateVariables.f90:28
constexpr double G_ROUNDOFF = 1.0e-24; // Cutcell_Mod.f90:45
constexpr double G_EDGE_EPS = 1.0e-14; // Edge_Geo_Cal's solid test
constexpr int MAX_CLOSE_PASSES = 4096;

// numpy's maximum / minimum (first operand on ties, NaN propagates)
__device__ inline double np_max(double a, double b) { return (a >= b || a != a) ? a : b; }
__device__ inline double np_min(double a, double b) { return (a <= b || a != a) ? a : b; }
__device__ inline double np_sign(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : x); }

struct Shape {
    int kind;
    double cx, cy, a, b, c;
};

__device__ void sdf_box(double X, double Y, double cx, double cy, double hw, double hh, double& phi, double& gx,
                        double& gy) {
    const double qx = fabs(X - cx) - hw;
    const double qy = fabs(Y - cy) - hh;
    const double ox = np_max(qx, 0.0);
    const double oy = np_max(qy, 0.0);
    const double outside = sqrt(ox * ox + oy * oy);
    const double inside = np_min(np_max(qx, qy), 0.0);
    phi = outside + inside;
    double sx = np_sign(X - cx);
    double sy = np_sign(Y - cy);
    if (sx == 0.0) sx = 1.0;
    if (sy == 0.0) sy = 1.0;
    if (outside > 0.0) {
        gx = sx * ox / outside;
        gy = sy * oy / outside;
    } else {
        gx = qx >= qy ? sx : 0.0;
        gy = qx >= qy ? 0.0 : sy;
    }
}

__device__ void sdf_circle(double X, double Y, double cx, double cy, double r, double& phi, double& gx,
                           double& gy) {
    const double rx = X - cx;
    const double ry = Y - cy;
    const double d = sqrt(rx * rx + ry * ry);
    phi = d - r;
    gx = d > 0.0 ? rx / d : 1.0;
    gy = d > 0.0 ? ry / d : 0.0;
}

// signed distance (> 0 in the fluid) and its unit gradient
__device__ void plane(const Shape& s, double X, double Y, double& phi, double& gx, double& gy) {
    if (s.kind == SG_SOLID_BOOM) {
        // union of the cylinder and the bar: the smaller distance (the
        // cylinder's on ties)
        double p1, x1, y1, p2, x2, y2;
        sdf_circle(X, Y, s.cx, s.cy, s.a, p1, x1, y1);
        const double bottom = s.cy - sqrt(s.a * s.a - s.b * s.b) - s.c;
        sdf_box(X, Y, s.cx, 0.5 * (s.cy + bottom), s.b, 0.5 * (s.cy - bottom), p2, x2, y2);
        const bool bar = p2 < p1;
        phi = bar ? p2 : p1;
        gx = bar ? x2 : x1;
        gy = bar ? y2 : y1;
    } else if (s.kind == SG_SOLID_BOX) {
        sdf_box(X, Y, s.cx, s.cy, s.a, s.b, phi, gx, gy);
    } else if (s.kind == SG_SOLID_CIRCLE) {
        sdf_circle(X, Y, s.cx, s.cy, s.a, phi, gx, gy);
    } else if (s.kind == SG_SOLID_ELLIPSE) {
        const double rx = (X - s.cx) / s.a;
        const double ry = (Y - s.cy) / s.b;
        const double F = sqrt(rx * rx + ry * ry);
        const double Fs = F > 0.0 ? F : 1.0;
        const double ex = rx / s.a / Fs;
        const double ey = ry / s.b / Fs;
        double gn = sqrt(ex * ex + ey * ey);
        if (!(gn > 0.0)) gn = 1.0;
        phi = (F - 1.0) / gn;
        gx = ex / gn;
        gy = ey / gn;
    } else {
        // no solid: a far plane
        phi = 1.0e3;
        gx = 0.0;
        gy = 1.0;
    }
}

// area fraction of a dx*dy cell centred on 0 where nx x + ny y + d < 0
__device__ double geo_frac(double nx, double ny, double d, double dx, double dy) {
    const double tnx = fabs(nx), tny = fabs(ny);
    const double half = 0.5 * (tnx * dx + tny * dy);
    if (d + half <= 0.0) return 1.0;
    if (d - half >= 0.0) return 0.0;
    if (tnx <= 1e-20) return 0.5 - d / dy;
    if (tny <= 1e-20) return 0.5 - d / dx;
    const double xx = (0.5 * tny * dy - d) / dx / tnx + 0.5;
    const double yy = (0.5 * tnx * dx - d) / dy / tny + 0.5;
    const double total = 0.5 * xx * yy;
    double top = 0.0, right = 0.0;
    if (yy > 1.0) {
        const double t = (yy - 1.0) / yy;
        top = t * t;
    }
    if (xx > 1.0) {
        const double t = (xx - 1.0) / xx;
        right = t * t;
    }
    return total * (1.0 - top - right);
}

// solid fraction of the edge between corner values l1, l2
__device__ double edge_solid(double l1, double l2) {
    const bool s1 = l1 < G_EDGE_EPS, s2 = l2 < G_EDGE_EPS;
    if (s1 && s2) return 1.0;
    double den = fabs(l1) + fabs(l2);
    if (!(den > 0.0)) den = 1.0;
    if (!s1 && s2) return fabs(l2) / den;
    if (s1 && !s2) return fabs(l1) / den;
    return 0.0;
}

struct GridSpec {
    int ni, nj;
    const double *xc, *yc, *dx, *dy;  // P centres and widths
    const double *hx, *hy;            // this grid's cell widths (per i, per j)
    int shift_x, shift_y;             // U: centres at x + dx/2; V: at y + dy/2
};

__device__ inline double grid_x(const GridSpec& G, int i) { return G.shift_x ? G.xc[i] + 0.5 * G.dx[i] : G.xc[i]; }
__device__ inline double grid_y(const GridSpec& G, int j) { return G.shift_y ? G.yc[j] + 0.5 * G.dy[j] : G.yc[j]; }

struct GeoOut {
    double *phi, *nx, *ny, *vofS, *wl;  // cell arrays
    double *aE, *aW, *aN, *aS;          // ghosted
    double *fcex, *fcey, *fcnx, *fcny;  // ghosted face centres FCE, FCN
};

__device__ inline size_t gix(int ni, int i, int j) { return (size_t)(i + 1) + (size_t)(j + 1) * (size_t)(ni + 2); }

// Edge_Geo_Cal (Cutcell_Mod.f90:302-337): solid fraction of the edge p1-p2
// and the centre of its solid part
__device__ void edge_geo(double p1x, double p1y, double p2x, double p2y, double nx, double ny, double d, double& a,
                         double& fx, double& fy) {
#pragma clang fp contract(off)
    const double l1 = nx * p1x + ny * p1y + d;
    const double l2 = nx * p2x + ny * p2y + d;
    fx = 0.5 * (p1x + p2x);
    fy = 0.5 * (p1y + p2y);
    a = 0.0;
    if (l1 >= G_EDGE_EPS && l2 >= G_EDGE_EPS) {
        a = 0.0;
    } else if (l1 < G_EDGE_EPS && l2 < G_EDGE_EPS) {
        a = 1.0;
    } else if (l1 >= G_EDGE_EPS && l2 < G_EDGE_EPS) {
        a = fabs(l2) / (fabs(l1) + fabs(l2));
        if (fabs(p1x - p2x) <= 1.0e-10)
            fy = p2y - 0.5 * a * (p2y - p1y);
        else
            fx = p2x - 0.5 * a * (p2x - p1x);
    } else if (l1 < G_EDGE_EPS && l2 >= G_EDGE_EPS) {
        a = fabs(l1) / (fabs(l1) + fabs(l2));
        if (fabs(p1x - p2x) <= 1.0e-10)
            fy = p1y - 0.5 * a * (p1y - p2y);
        else
            fx = p1x - 0.5 * a * (p1x - p2x);
    }
}

// the four edges of a cell of size hx*hy with solid line (nx, ny, phi):
// solid fractions and raw centres of their solid parts (Cutcell_Mod.f90:52-73)
struct Edges {
    double sW, sN, sE, sS;
    double wx, wy, nx, ny, ex, ey, sx, sy;
};

__device__ Edges cell_edges(double nx, double ny, double phi, double hx, double hy) {
#pragma clang fp contract(off)
    const double X0 = -(hx * 0.5), X1 = -(hx * -0.5);
    const double Y0 = -(hy * 0.5), Y1 = -(hy * -0.5);
    Edges e;
    edge_geo(X0, Y0, X0, Y1, nx, ny, phi, e.sW, e.wx, e.wy);
    edge_geo(X0, Y1, X1, Y1, nx, ny, phi, e.sN, e.nx, e.ny);
    edge_geo(X1, Y1, X1, Y0, nx, ny, phi, e.sE, e.ex, e.ey);
    edge_geo(X1, Y0, X0, Y0, nx, ny, phi, e.sS, e.sx, e.sy);
    return e;
}

// fluid face centre from the raw centre c of the solid part and the fluid
// fraction a (Cutcell_Mod.f90:124-158)
__device__ inline double fluid_centre(double c, double a, double h) {
#pragma clang fp contract(off)
    double o = fabs(c) < EPSI ? 0.0 : (0.5 - a / 2.0) * copysign(1.0, -c) * h;
    return a < G_EPSIF ? 0.0 : o;
}

// level set, raw solid fraction and raw wall length of every cell
// (SolidVolumeFraction's role, Clsvof_Mod.f90:273-277; Cutcell_Mod.f90:79-89)
__global__ void __launch_bounds__(256) k_geo_cells(GridSpec G, Shape S, GeoOut o) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= G.ni || j >= G.nj) return;
    const double X = grid_x(G, i), Y = grid_y(G, j);
    const double hx = G.hx[i], hy = G.hy[j];
    double phi, nx, ny;
    plane(S, X, Y, phi, nx, ny);
    double v = geo_frac(nx, ny, phi, hx, hy);
    if (v < G_VOFEPS) v = 0.0;
    if (v >= 1.0 - G_VOFEPS) v = 1.0;
    const Edges e = cell_edges(nx, ny, phi, hx, hy);
    const double wx = (e.sE - e.sW) * hx, wy = (e.sN - e.sS) * hy;
    const size_t c = (size_t)i + (size_t)j * G.ni;
    o.phi[c] = phi;
    o.nx[c] = nx;
    o.ny[c] = ny;
    o.vofS[c] = v;
    o.wl[c] = sqrt(wx * wx + wy * wy);
}

__device__ inline Edges edges_of(const GridSpec& G, const GeoOut& o, int i, int j) {
    const size_t c = (size_t)i + (size_t)j * G.ni;
    return cell_edges(o.nx[c], o.ny[c], o.phi[c], G.hx[i], G.hy[j]);
}

// FCE of the ghost-filling stage of Cell_Geo_Cal (raw centres; columns 0
// and ni+1 shifted by the first / last width, Cutcell_Mod.f90:91-99) at
// ghosted column gi of cell row j
__device__ inline void raw_fce(const GridSpec& G, const GeoOut& o, int gi, int j, double& x, double& y) {
#pragma clang fp contract(off)
    const int ni = G.ni;
    const int i = gi < 1 ? 0 : (gi > ni ? ni - 1 : gi - 1);
    const Edges e = edges_of(G, o, i, j);
    x = e.ex;
    y = e.ey;
    if (gi == 0) x = e.ex - G.hx[0];
    if (gi == ni + 1) x = e.ex + G.hx[ni - 1];
}

// apertures and face centres after the conversion to fluid values and the
// harmonic averaging of i = 2..Isize-1, j = 2..Jsize-1 (Cutcell_Mod.f90:
// 91-204); one thread per ghosted entry, each evaluating the edges of its
// cell and of the neighbours sharing its faces.  Ghost apertures are 0, ghost
// face centres follow :91-117 (the rows of FCN copy FCE).
__global__ void __launch_bounds__(256) k_geo_faces(GridSpec G, GeoOut o) {
#pragma clang fp contract(off)
    const int ni = G.ni, nj = G.nj;
    const int gi = blockIdx.x * TX + threadIdx.x;
    const int gj = blockIdx.y * TY + threadIdx.y;
    if (gi >= ni + 2 || gj >= nj + 2) return;
    const size_t g = (size_t)gi + (size_t)gj * (size_t)(ni + 2);
    const int i = gi - 1, j = gj - 1;
    if (i < 0 || i >= ni || j < 0 || j >= nj) {
        o.aE[g] = 0.0;
        o.aW[g] = 0.0;
        o.aN[g] = 0.0;
        o.aS[g] = 0.0;
        if (gj >= 1 && gj <= nj) {
            // columns 0 and ni+1 of rows 1..nj
            const Edges e = edges_of(G, o, gi == 0 ? 0 : ni - 1, j);
            const double sh = gi == 0 ? -G.hx[0] : G.hx[ni - 1];
            o.fcex[g] = e.ex + sh;
            o.fcey[g] = e.ey;
            o.fcnx[g] = e.nx + sh;
            o.fcny[g] = e.ny;
        } else {
            // rows 0 and nj+1 (every column) from FCE of row 1 / nj
            double x, y;
            raw_fce(G, o, gi, gj == 0 ? 0 : nj - 1, x, y);
            const double yy = gj == 0 ? y - G.hy[0] : y + G.hy[nj - 1];
            o.fcex[g] = x;
            o.fcey[g] = yy;
            o.fcnx[g] = x;
            o.fcny[g] = yy;
        }
        return;
    }
    const double hx = G.hx[i], hy = G.hy[j];
    const Edges e = edges_of(G, o, i, j);
    const double aEc = 1.0 - e.sE, aWc = 1.0 - e.sW, aNc = 1.0 - e.sN, aSc = 1.0 - e.sS;
    const double eyc = fluid_centre(e.ey, aEc, hy);
    const double nxc = fluid_centre(e.nx, aNc, hx);
    const bool avg = ni >= 3 && nj >= 3;
    const bool in_i = i >= 1 && i <= ni - 2, in_j = j >= 1 && j <= nj - 2;
    double aE = aEc, fcex = e.ex, fcey = eyc;
    if (avg && in_i && in_j) {
        const Edges b = edges_of(G, o, i + 1, j);
        const double w_ = 1.0 - b.sW, e_ = aEc;
        const double wyc = fluid_centre(b.wy, w_, hy);
        aE = 2.0 * e_ * w_ / (e_ + w_ + G_ROUNDOFF);
        fcex = (e.ex * e_ + (b.wx + (grid_x(G, i + 1) - grid_x(G, i))) * w_) / (e_ + w_ + G_ROUNDOFF);
        fcey = (eyc * e_ + wyc * w_) / (e_ + w_ + G_ROUNDOFF);
        if (e_ < G_EPSIF && w_ < G_EPSIF) fcex = hx / 2.0;
    }
    double aN = aNc, fcnx = nxc, fcny = e.ny;
    if (avg && in_i && in_j) {
        const Edges b = edges_of(G, o, i, j + 1);
        const double s_ = 1.0 - b.sS, n_ = aNc;
        const double sxc = fluid_centre(b.sx, s_, hx);
        aN = 2.0 * n_ * s_ / (n_ + s_ + G_ROUNDOFF);
        fcnx = (nxc * n_ + sxc * s_) / (n_ + s_ + G_ROUNDOFF);
        fcny = (e.ny * n_ + (b.sy + (grid_y(G, j + 1) - grid_y(G, j))) * s_) / (n_ + s_ + G_ROUNDOFF);
        if (n_ < G_EPSIF && s_ < G_EPSIF) fcny = hy / 2.0;
    }
    double aW = aWc;
    if (avg && i - 1 >= 1 && i - 1 <= ni - 2 && in_j) {
        const double e_ = 1.0 - edges_of(G, o, i - 1, j).sE;
        aW = 2.0 * e_ * aWc / (e_ + aWc + G_ROUNDOFF);
    }
    double aS = aSc;
    if (avg && in_i && j - 1 >= 1 && j - 1 <= nj - 2) {
        const double n_ = 1.0 - edges_of(G, o, i, j - 1).sN;
        aS = 2.0 * n_ * aSc / (n_ + aSc + G_ROUNDOFF);
    }
    o.aE[g] = aE;
    o.aW[g] = aW;
    o.aN[g] = aN;
    o.aS[g] = aS;
    o.fcex[g] = fcex;
    o.fcey[g] = fcey;
    o.fcnx[g] = fcnx;
    o.fcny[g] = fcny;
}

__device__ inline void close_faces(int ni, size_t g, const GeoOut& o) {
    o.aW[g] = 0.0;
    o.aE[g] = 0.0;
    o.aN[g] = 0.0;
    o.aS[g] = 0.0;
    o.aE[g - 1] = 0.0;
    o.aN[g - (size_t)(ni + 2)] = 0.0;
    o.aW[g + 1] = 0.0;
    o.aS[g + (size_t)(ni + 2)] = 0.0;
}

// one pass of the closing loop (Cutcell_Mod.f90:214-243).  The reference
// visits i outer, j inner: when (i, j) is visited its west and south faces
// carry the closings of (i-1, j) and (i, j-1), its east and north faces
// are still the averaged values (the cells that close them come later).  A
// cell closes if it is solid, or (P cells) if with those faces it is shut
// on both sides along x or along y.  Iterated from no closings this reaches
// the sequential result (the update is monotone over an acyclic graph).
__global__ void __launch_bounds__(256) k_geo_close(int ni, int nj, int p_cell, GeoOut o, const int* __restrict__ dold,
                                                   int* __restrict__ dnew, int* __restrict__ changed) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    int d = o.vofS[c] > 1.0 - EPSI ? 1 : 0;
    if (!d && p_cell) {
        const double w = (i > 0 && dold[c - 1]) ? 0.0 : o.aW[g];
        const double s = (j > 0 && dold[c - ni]) ? 0.0 : o.aS[g];
        d = (w + o.aE[g] < G_ROUNDOFF || s + o.aN[g] < G_ROUNDOFF) ? 1 : 0;
    }
    dnew[c] = d;
    if (d != dold[c]) atomicOr(changed, 1);
}

// closed cells become solid and shut their faces and their neighbours'
// (the writes are zeros only)
__global__ void __launch_bounds__(256) k_geo_close_apply(int ni, int nj, GeoOut o, const int* __restrict__ d) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni;
    if (d[c]) {
        o.vofS[c] = 1.0;
        close_faces(ni, gix(ni, i, j), o);
    }
}

// wall length of full solid cells (Cutcell_Mod.f90:245-258)
__global__ void __launch_bounds__(256) k_geo_wall(GridSpec G, GeoOut o) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= G.ni || j >= G.nj) return;
    const size_t c = (size_t)i + (size_t)j * G.ni, g = gix(G.ni, i, j);
    if (!(fabs(o.vofS[c] - 1.0) < G_ROUNDOFF)) return;
    if (o.aW[g] == 0.0 || o.aE[g] == 0.0) o.wl[c] = G.hy[j];
    if (o.aN[g] == 0.0 || o.aS[g] == 0.0) o.wl[c] = G.hx[i];
}

// corner patch of the P cells (ICorProb, Cutcell_Mod.f90:259-286) for
// j = 1..Jsize-4: a fluid cell under two cells 60 % solid (the boom's bar)
// or under a full solid cell.  The reference's north aperture is the
// default-real literal 0.40 and the south one of the cell above 0.4d0;
// both are 0.4 here so the face keeps one aperture (symmetric rows).
__global__ void __launch_bounds__(256) k_geo_corner(GridSpec G, GeoOut o) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int ni = G.ni;
    if (i >= ni || j >= G.nj - 4) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const double s = o.vofS[c];
    if (!(s > -EPSI && s < EPSI)) return;
    const double s1 = o.vofS[c + ni], s2 = o.vofS[c + 2 * (size_t)ni];
    if (s1 > 0.6 - EPSI && s1 < 0.6 + EPSI && s2 > 0.6 - EPSI && s2 < 0.6 + EPSI) {
        o.aN[g] = 0.4;
        o.aS[g + (size_t)(ni + 2)] = 0.4;
        o.fcnx[g] = -0.3 * G.hx[i];
        o.fcny[g] = 0.5 * G.hy[j];
        o.wl[c] = 0.6 * G.hx[i];
        o.ny[c] = -1.0;
        o.nx[c] = 0.0;
    }
    if (s1 > 1.0 - EPSI && s1 < 1.0 + EPSI) {
        o.wl[c] = G.hx[i];
        o.ny[c] = -1.0;
        o.nx[c] = 0.0;
    }
}

struct MoexIn {
    int ni, nj;
    const double *vofS, *vofSU, *vofSV;
    double *aE, *aW, *aN, *aS;
    int *moexU, *moexV, *uflag, *vflag;
};

// DefineMomentumExchangeCell (Cutcell_Mod.f90:360-397) and the U/V flags
// Posnu /= -1 .and. MoExCell /= 1; the U column i = Isize and the V row
// j = Jsize are never numbered by NumberExternalCell and keep Posnu = 0
__global__ void __launch_bounds__(256) k_geo_moex(MoexIn m) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int ni = m.ni, nj = m.nj;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const double su = m.vofSU[c], sv = m.vofSV[c];
    int mu = 0, mv = 0;
    if (su >= EPSI && su < 1.0 - EPSI) {
        if (i < ni - 1) {
            if (m.vofS[c] >= 1.0 - EPSI || m.vofS[c + 1] >= 1.0 - EPSI || m.aE[g] < G_EPSIF) {
                mu = 1;
                m.aE[g] = 0.0;
                m.aW[g + 1] = 0.0;
            }
        } else if (m.vofS[c] >= 1.0 - EPSI) {
            mu = 1;
        }
    }
    if (sv >= EPSI && sv < 1.0 - EPSI) {
        if (j < nj - 1) {
            if (m.vofS[c] >= 1.0 - EPSI || m.vofS[c + ni] >= 1.0 - EPSI || m.aN[g] < G_EPSIF) {
                mv = 1;
                m.aN[g] = 0.0;
                m.aS[g + (size_t)(ni + 2)] = 0.0;
            }
        } else if (m.vofS[c] >= 1.0 - EPSI) {
            mv = 1;
        }
    }
    m.moexU[c] = mu;
    m.moexV[c] = mv;
    const bool nu = su < 1.0 - EPSI || i == ni - 1;
    const bool nv = sv < 1.0 - EPSI || j == nj - 1;
    m.uflag[c] = (nu && mu != 1) ? 1 : 0;
    m.vflag[c] = (nv && mv != 1) ? 1 : 0;
}

struct StagIn {
    int ni, nj, which;   // 0: U cells, 1: V cells
    const double *hx, *hy;
    const double *sxn, *syn;  // BGrid spacing of the interpolation routines: per i (east), per j (north)
    const double *pdx, *pdy;  // P-cell widths
    const double *pfx, *pfy;  // the P cells' FCE (U) or FCN (V) after Cell_Geo_Cal (ghosted)
    const int* moex;
    double* a[STG_COUNT];
    int* act;
    int* moex_out;
};

__device__ inline bool stag_numbered(const StagIn& s, int i, int j) {
    return s.act[(size_t)i + (size_t)j * s.ni] != 0 || (s.which == 0 && i == s.ni - 1) ||
           (s.which == 1 && j == s.nj - 1);
}

// NewCellFace per cell (Cutcell_Mod.f90:399-543): the centroid offset from
// the P cells' face centres, or for a small cell next to one P cell from its
// own open faces and solid line; the wall distance; the wall length from
// the final apertures; the momentum-system row
__global__ void __launch_bounds__(256) k_stag_cells(StagIn s) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const size_t gs = g - (size_t)(ni + 2);
    const double v = s.a[STG_VOFS][c], nx = s.a[STG_NXS][c], ny = s.a[STG_NYS][c], phi = s.a[STG_PHIS][c];
    const int mo = s.moex[c];
    const double tol = 1.0e-24;
    double cx = 0.0, cy = 0.0;
    if (mo != 1 && v > EPSI) {
        if (s.which == 0) {
            cx = s.pfx[g] - s.pdx[i] / 2.0;
            cy = s.pfy[g];
        } else {
            cx = s.pfx[g];
            cy = s.pfy[g] - s.pdy[j] / 2.0;
        }
    }
    if (mo == 1 && v < 1.0 - EPSI) {
        const double nN = s.a[STG_AN][g], nS = s.a[STG_AN][gs];
        const double eE = s.a[STG_AE][g], eW = s.a[STG_AE][g - 1];
        const double cy1 = s.a[STG_FCEY][g] * eE / (eE + tol) + s.a[STG_FCEY][g - 1] * eW / (eW + tol);
        if (nN * nS == 0.0) {
            cx = s.a[STG_FCNX][g] * nN / (nN + tol) + s.a[STG_FCNX][gs] * nS / (nS + tol);
        } else {
            cy = cy1;
            cx = -(ny * cy + phi) / (nx + tol);
        }
        if (eE * eW == 0.0)
            cy = cy1;
        else
            cy = -(nx * cx + phi) / (ny + tol);
    }
    s.a[STG_CCX][c] = cx;
    s.a[STG_CCY][c] = cy;
    const double dl = fabs(cx * nx + cy * ny + phi) + tol;
    s.a[STG_DELH][c] = mo == 1 ? 1.0e-20 : dl;
    const double wx = (s.a[STG_AE][g] - s.a[STG_AW][g]) * s.hx[i];
    const double wy = (s.a[STG_AN][g] - s.a[STG_AS][g]) * s.hy[j];
    s.a[STG_WLLH][c] = sqrt(wx * wx + wy * wy);
    int active = v < 1.0 - EPSI ? 1 : 0;
    if ((s.which == 0 && i == ni - 1) || (s.which == 1 && j == nj - 1)) active = 0;
    s.act[c] = active;
    s.moex_out[c] = mo;
}

// EastFaceInterpolationInf (Cutcell_Mod.f90:551-616) for the pair of cells
// (i, j), (i+1, j), i <= ni-2
__device__ void east_pair(const StagIn& s, int i, int j, double& SxE, double& EtaE, double& AlE, double& DAlE) {
#pragma clang fp contract(off)
    const int ni = s.ni;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const double bd = s.sxn[i];
    if (!(stag_numbered(s, i, j) && stag_numbered(s, i + 1, j))) {
        AlE = 1.0;
        DAlE = 0.0;
        SxE = bd;
        EtaE = 0.5;
        return;
    }
    const double *ccx = s.a[STG_CCX], *ccy = s.a[STG_CCY];
    const double *nxS = s.a[STG_NXS], *nyS = s.a[STG_NYS], *phiS = s.a[STG_PHIS];
    const double fex = s.a[STG_FCEX][g], fey = s.a[STG_FCEY][g];
    SxE = ccx[c + 1] + bd - ccx[c];
    const double Sy = ccy[c + 1] - ccy[c];
    EtaE = fabs(fex - ccx[c]) / SxE;
    if (EtaE >= 1.0) EtaE = 0.5;
    const double xf = (1.0 - EtaE) * ccx[c] + EtaE * (ccx[c + 1] + bd);
    const double yf = (1.0 - EtaE) * ccy[c] + EtaE * ccy[c + 1];
    if (fabs(Sy) > 0.0001 * s.hy[j]) {
        const double nyf = 0.5 * (nyS[c] + nyS[c + 1]);
        const double ipd = 0.5 * (fabs(xf * nxS[c] + yf * nyS[c] + phiS[c]) +
                                  fabs((xf - bd) * nxS[c + 1] + yf * nyS[c + 1] + phiS[c + 1]));
        const double fcd = 0.5 * (fabs(fex * nxS[c] + fey * nyS[c] + phiS[c]) +
                                  fabs((fex - bd) * nxS[c + 1] + fey * nyS[c + 1] + phiS[c + 1]));
        AlE = fabs(fcd) / fabs(ipd);
        DAlE = Sy * nyf / (SxE * fabs(ipd));
        if (fabs(ipd) < 1.0e-5) {
            DAlE = 0.0;
            AlE = 1.0;
        }
        if (AlE >= 2.0) AlE = 2.0;
    } else {
        DAlE = 0.0;
        AlE = 1.0;
    }
}

// NorthFaceInterpolationInf (Cutcell_Mod.f90:618-700) for the pair (i, j),
// (i, j+1), j <= nj-2; its factor is IntPointDist / FaceCenterDist
__device__ void north_pair(const StagIn& s, int i, int j, double& SyN, double& EtaN, double& AlN, double& DAlN) {
#pragma clang fp contract(off)
    const int ni = s.ni;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j), u = c + (size_t)ni;
    const double bd = s.syn[j];
    if (!(stag_numbered(s, i, j) && stag_numbered(s, i, j + 1))) {
        EtaN = 0.5;
        AlN = 1.0;
        DAlN = 0.0;
        SyN = bd;
        return;
    }
    const double *ccx = s.a[STG_CCX], *ccy = s.a[STG_CCY];
    const double *nxS = s.a[STG_NXS], *nyS = s.a[STG_NYS], *phiS = s.a[STG_PHIS];
    const double fnx = s.a[STG_FCNX][g], fny = s.a[STG_FCNY][g];
    SyN = ccy[u] + bd - ccy[c];
    const double Sx = ccx[u] - ccx[c];
    EtaN = fabs(fny - ccy[c]) / SyN;
    if (EtaN >= 1.0) EtaN = 0.5;
    const double xf = (1.0 - EtaN) * ccx[c] + EtaN * ccx[u];
    const double yf = (1.0 - EtaN) * ccy[c] + EtaN * (ccy[u] + bd);
    if (fabs(Sx) >= 0.0001 * s.hx[i]) {
        const double nxf = 0.5 * (nxS[c] + nxS[u]);
        const double ipd = 0.5 * (fabs(xf * nxS[c] + yf * nyS[c] + phiS[c]) +
                                  fabs(xf * nxS[u] + (yf - bd) * nyS[u] + phiS[u]));
        const double fcd = 0.5 * (fabs(fnx * nxS[c] + fny * nyS[c] + phiS[c]) +
                                  fabs(fnx * nxS[u] + (fny - bd) * nyS[u] + phiS[u]));
        AlN = fabs(ipd) / fabs(fcd);
        DAlN = Sx * nxf / (SyN * fabs(ipd));
        if (fabs(ipd) < 1.0e-5) {
            DAlN = 0.0;
            AlN = 1.0;
        }
        if (AlN >= 2.0) AlN = 2.0;
    } else {
        DAlN = 0.0;
        AlN = 1.0;
    }
}

// east factors of every ghosted entry: the pair of its column (columns 0,
// ni and ni+1 copy their nearest pair, rows 0 and nj+1 their nearest row);
// AlE of column 0 is AlN(1, j) as the previous geometry left it
// (Cutcell_Mod.f90:606: this call's north factors are not formed yet)
__global__ void __launch_bounds__(256) k_stag_east(StagIn s) {
    const int gi = blockIdx.x * TX + threadIdx.x;
    const int gj = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (gi >= ni + 2 || gj >= nj + 2) return;
    const int j = (gj < 1 ? 1 : (gj > nj ? nj : gj)) - 1;
    const int pi = gi <= 1 ? 0 : (gi >= ni ? ni - 2 : gi - 1);
    double SxE, EtaE, AlE, DAlE;
    east_pair(s, pi, j, SxE, EtaE, AlE, DAlE);
    if (gi == 0) AlE = s.a[STG_ALN][gix(ni, 0, j)];
    const size_t g = (size_t)gi + (size_t)gj * (size_t)(ni + 2);
    s.a[STG_SXE][g] = SxE;
    s.a[STG_ETAE][g] = EtaE;
    s.a[STG_ALE][g] = AlE;
    s.a[STG_DALE][g] = DAlE;
}

// north factors of every ghosted entry: the pair of its row (rows 0, nj and
// nj+1 copy their nearest pair, columns 0 and ni+1 their nearest column)
__global__ void __launch_bounds__(256) k_stag_north(StagIn s) {
    const int gi = blockIdx.x * TX + threadIdx.x;
    const int gj = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (gi >= ni + 2 || gj >= nj + 2) return;
    const int i = (gi < 1 ? 1 : (gi > ni ? ni : gi)) - 1;
    const int pj = gj <= 1 ? 0 : (gj >= nj ? nj - 2 : gj - 1);
    double SyN, EtaN, AlN, DAlN;
    north_pair(s, i, pj, SyN, EtaN, AlN, DAlN);
    const size_t g = (size_t)gi + (size_t)gj * (size_t)(ni + 2);
    s.a[STG_SYN][g] = SyN;
    s.a[STG_ETAN][g] = EtaN;
    s.a[STG_ALN][g] = AlN;
    s.a[STG_DALN][g] = DAlN;
}

struct ForceIn {
    int ni, nj;
    const double *dx, *vofS, *nyS, *wl, *aE, *aW, *p;  // P cells (p, aE, aW ghosted)
    const double *vofSV, *delhV, *v;                  // V cells (v ghosted)
    const int* moexV;
    double vs, Rey;
};

constexpr int NTF_FORCE = 256;

// per-block partial sums of the pressure (clp) and shear (clf) parts of
// ComputeForceObject (Clsvof_Mod.f90:2400-2549, ICorProb = .TRUE.)
__global__ void __launch_bounds__(NTF_FORCE) k_force(ForceIn f, double* part, int nb) {
#pragma clang fp contract(off)
    __shared__ double red[2][NTF_FORCE / 64];
    const int ni = f.ni, nj = f.nj;
    double clp = 0.0, clf = 0.0;
    for (int e = blockIdx.x * NTF_FORCE + threadIdx.x; e < ni * nj; e += gridDim.x * NTF_FORCE) {
        const int i = e % ni, j = e / ni;
        const size_t c = (size_t)e, g = gix(ni, i, j);
        const double s = f.vofS[c];
        if (s < 1.0 - EPSI && s > EPSI) clp += f.p[g] * f.wl[c] * f.nyS[c];
        if (j <= nj - 2) {
            const double sn = f.vofS[c + ni];
            const bool full_above = sn < 1.0 + EPSI && sn > 1.0 - EPSI;
            if (s < 1.0 - EPSI && f.aE[g] < G_EPSIF && f.aW[g] < G_EPSIF && full_above) {
                const double pres = f.p[g - (size_t)(ni + 2)], area = f.dx[i], nyfp = -1.0;
                clp += pres * area * nyfp - f.p[g] * area * nyfp;
            }
            if (s > -EPSI && s < EPSI) {
                if (full_above) clp += f.p[g] * f.dx[i] * -1.0;
                if (sn > 0.6 - EPSI && sn < 0.6 + EPSI && j < nj - 2) {
                    const double s2 = f.vofS[c + 2 * (size_t)ni];
                    if (s2 > 0.6 - EPSI && s2 < 0.6 + EPSI) clp += f.p[g] * (0.6 * f.dx[i]) * -1.0;
                }
            }
        }
        const double sv = f.vofSV[c];
        if (sv > EPSI && sv < 1.0 - EPSI && f.moexV[c] != 1) clf += (f.v[g] - f.vs) / f.Rey / f.delhV[c] * f.wl[c];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        clp += __shfl_xor(clp, o, 64);
        clf += __shfl_xor(clf, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = clp;
        red[1][threadIdx.x >> 6] = clf;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double r = red[threadIdx.x][0];
        for (int w = 1; w < NTF_FORCE / 64; ++w) r += red[threadIdx.x][w];
        part[(size_t)threadIdx.x * nb + blockIdx.x] = r;
    }
}

// one block: the partials in block order; out = {clf - clp, clp, clf}
__global__ void __launch_bounds__(NTF_FORCE) k_force_fin(const double* part, int nb, double* out) {
    __shared__ double red[2][NTF_FORCE / 64];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nb; k += NTF_FORCE) {
        a += part[k];
        b += part[(size_t)nb + k];
    }
    for (int o = 32; o >= 1; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a;
        red[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double clp = red[0][0], clf = red[1][0];
        for (int w = 1; w < NTF_FORCE / 64; ++w) {
            clp += red[0][w];
            clf += red[1][w];
        }
        out[0] = clf - clp;
        out[1] = clp;
        out[2] = clf;
    }
}

// SolidVolumeFraction (Clsvof_Mod.f90:2553-2929) on the P cells for a shape:
// the solid level set, normal and fraction (snapped below epsi and above
// 1 - epsi, :2906-2907), and the liquid fraction carried onto the new fluid
// part of the cell (:2908-2917)
__global__ void __launch_bounds__(256) k_solid_fraction(GridSpec G, Shape S, double* __restrict__ phiS,
                                                        double* __restrict__ nxS, double* __restrict__ nyS,
                                                        double* __restrict__ vofS, double* __restrict__ vof) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= G.ni || j >= G.nj) return;
    const size_t c = (size_t)i + (size_t)j * G.ni;
    double phi, nx, ny;
    plane(S, G.xc[i], G.yc[j], phi, nx, ny);
    double vs = geo_frac(nx, ny, phi, G.hx[i], G.hy[j]);
    if (vs < EPSI) vs = 0.0;
    if (vs >= 1.0 - EPSI) vs = 1.0;
    const double old = vofS[c];
    double f = vof[c];
    if (f + old >= 1.0 - EPSI || f + vs >= 1.0 - EPSI) f = 1.0 - vs;
    if (old < 1.0 - EPSI && f + old < 1.0 - EPSI && f + vs < 1.0 - EPSI) f = f / (1.0 - old + G_ROUNDOFF) * (1.0 - vs);
    phiS[c] = phi;
    nxS[c] = nx;
    nyS[c] = ny;
    vofS[c] = vs;
    vof[c] = f;
}

struct NewVarIn {
    int ni, nj;
    const double *vofSUo, *vofSU, *vofSVo, *vofSV;  // cell arrays
    const double *aEo, *aE, *aNo, *aN;              // ghosted P apertures
    double vb;
    double *u, *v;                                   // ghosted
};

// InterNewVar (Solver_Mod.f90:360-495): every (i,j) reads and writes only
// its own u and v (the reference's uleft/uright/vbot/vtop and mass-cell
// lookups are computed but unused), so the cells are independent
__global__ void __launch_bounds__(256) k_inter_new_var(NewVarIn a) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= a.ni || j >= a.nj) return;
    const size_t c = (size_t)i + (size_t)j * a.ni, g = gix(a.ni, i, j);
    double u = a.u[g], v = a.v[g];
    if (a.vofSUo[c] > 1.0 - EPSI && a.vofSU[c] < 1.0 - EPSI) u = 0.0;
    if (a.vofSVo[c] > 1.0 - EPSI && a.vofSV[c] < 1.0 - EPSI) v = a.vb;
    const double nN = a.aN[g], oN = a.aNo[g];
    if (nN > EPSI && oN > EPSI && (nN > 0.5 || oN > 0.5) && oN / nN < 0.6) v = v * oN / nN;
    const double nE = a.aE[g], oE = a.aEo[g];
    if (nE > EPSI && oE > EPSI && (nE > 0.5 || oE > 0.5) && oE / nE < 0.6) u = u * oE / nE;
    a.u[g] = u;
    a.v[g] = v;
}

}  // namespace

int solid_fraction_step(Ctx* c, const sg_solid_shape* shape) {
    const GridSpec GP{c->ni, c->nj, c->cls.xc, c->cls.yc, c->dx, c->dy, c->dx, c->dy, 0, 0};
    const Shape S{shape->kind, shape->cx, shape->cy, shape->a, shape->b, shape->c};
    auto F = [&](int f) { return (double*)c->field[f]; };
    hipLaunchKernelGGL(k_solid_fraction, grid2d(c->ni, c->nj), block2d(), 0, c->stream, GP, S, F(SG_F_PHIS),
                       F(SG_F_NXS), F(SG_F_NYS), F(SG_F_VOFS), F(SG_F_VOF));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int geometry_snapshot(Ctx* c) {
    GeoScratch& W = c->geo;
    const size_t ncell = (size_t)c->ni * c->nj, nghost = (size_t)(c->ni + 2) * (c->nj + 2);
    if (!W.old[0]) {
        W.old[0] = (double*)dev_alloc(c, nghost * sizeof(double));
        W.old[1] = (double*)dev_alloc(c, nghost * sizeof(double));
        W.old[2] = (double*)dev_alloc(c, ncell * sizeof(double));
        W.old[3] = (double*)dev_alloc(c, ncell * sizeof(double));
        if (!W.old[0] || !W.old[1] || !W.old[2] || !W.old[3]) {
            set_error("sg_geometry_snapshot: device allocation failed");
            return SG_ERR_ALLOC;
        }
    }
    const int src[4] = {SG_F_AE, SG_F_AN, SG_F_VOFSU, SG_F_VOFSV};
    for (int k = 0; k < 4; ++k)
        SG_HIP(hipMemcpyAsync(W.old[k], c->field[src[k]], c->field_bytes[src[k]], hipMemcpyDeviceToDevice, c->stream));
    W.have_old = true;
    return SG_OK;
}

int inter_new_var(Ctx* c, double vb) {
    GeoScratch& W = c->geo;
    if (!W.have_old) {
        set_error("sg_inter_new_var: no geometry snapshot (call sg_geometry_snapshot at the start of the step)");
        return SG_ERR_STATE;
    }
    auto F = [&](int f) { return (double*)c->field[f]; };
    const NewVarIn a{c->ni,     c->nj,        W.old[2],       F(SG_F_VOFSU), W.old[3], F(SG_F_VOFSV),
                     W.old[0],  F(SG_F_AE),   W.old[1],       F(SG_F_AN),    vb,       F(SG_F_U),
                     F(SG_F_V)};
    hipLaunchKernelGGL(k_inter_new_var, grid2d(c->ni, c->nj), block2d(), 0, c->stream, a);
    SG_HIP(hipGetLastError());
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

int object_force(Ctx* c, double vs, double* force, double* clp, double* clf) {
    const int ni = c->ni, nj = c->nj;
    auto F = [&](int f) { return (const double*)c->field[f]; };
    const StagDev& V = c->stag[1];
    const ForceIn in{ni,        nj,        c->dx,       F(SG_F_VOFS),      F(SG_F_NYS),   F(SG_F_WLLH),
                     F(SG_F_AE), F(SG_F_AW), F(SG_F_P),   V.a[STG_VOFS],     V.a[STG_DELH], F(SG_F_V),
                     V.moex,    vs,        c->prm.Re};
    int nb = (ni * nj + NTF_FORCE - 1) / NTF_FORCE;
    if (nb > DIAG_BLOCKS) nb = DIAG_BLOCKS;
    // the diagnostics' partial buffer (same stream, so never in use by both)
    double* part = c->diag.partials;
    double* out = c->step_dev + 1;
    hipLaunchKernelGGL(k_force, dim3(nb), dim3(NTF_FORCE), 0, c->stream, in, part, nb);
    hipLaunchKernelGGL(k_force_fin, dim3(1), dim3(NTF_FORCE), 0, c->stream, (const double*)part, nb, out);
    SG_HIP(hipGetLastError());
    double h[3];
    SG_HIP(hipMemcpyAsync(h, out, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    if (force) *force = h[0];
    if (clp) *clp = h[1];
    if (clf) *clf = h[2];
    return SG_OK;
}

int geometry_rebuild(Ctx* c, const sg_solid_shape* shape) {
    const int ni = c->ni, nj = c->nj;
    const size_t ncell = (size_t)ni * nj, nghost = (size_t)(ni + 2) * (nj + 2);
    GeoScratch& W = c->geo;
    if (!W.pfc[0]) {
        for (double*& p : W.pfc) p = (double*)dev_alloc(c, nghost * sizeof(double));
        for (int*& p : W.close) p = (int*)dev_alloc(c, ncell * sizeof(int));
        W.changed = (int*)dev_alloc(c, sizeof(int));
        W.moex[0] = (int*)dev_alloc(c, ncell * sizeof(int));
        W.moex[1] = (int*)dev_alloc(c, ncell * sizeof(int));
        if (!W.pfc[3] || !W.close[1] || !W.changed || !W.moex[0] || !W.moex[1]) {
            set_error("sg_geometry_rebuild: device allocation failed");
            return SG_ERR_ALLOC;
        }
    }
    hipStream_t st = c->stream;
    const Shape S{shape->kind, shape->cx, shape->cy, shape->a, shape->b, shape->c};
    const dim3 gc = grid2d(ni, nj), gg = grid2d(ni + 2, nj + 2), b = block2d();
    auto F = [&](int f) { return (double*)c->field[f]; };

    // the closing loop: solid cells, and on the P cells the cells shut on
    // both sides, repeated until no decision changes
    auto close_cells = [&](const GeoOut& O, int p_cell) -> int {
        SG_HIP(hipMemsetAsync(W.close[0], 0, ncell * sizeof(int), st));
        int cur = 0, passes = 0;
        for (;;) {
            int changed = 0;
            SG_HIP(hipMemsetAsync(W.changed, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_geo_close, gc, b, 0, st, ni, nj, p_cell, O, (const int*)W.close[cur],
                               W.close[1 - cur], W.changed);
            SG_HIP(hipGetLastError());
            cur = 1 - cur;
            if (!p_cell) break;  // the solid test alone: one pass
            SG_HIP(hipMemcpyAsync(&changed, W.changed, sizeof(int), hipMemcpyDeviceToHost, st));
            SG_HIP(hipStreamSynchronize(st));
            if (!changed) break;
            if (++passes > MAX_CLOSE_PASSES) {
                set_error("sg_geometry_rebuild: the closing rule did not settle");
                return SG_ERR_STATE;
            }
        }
        hipLaunchKernelGGL(k_geo_close_apply, gc, b, 0, st, ni, nj, O, (const int*)W.close[cur]);
        SG_HIP(hipGetLastError());
        return SG_OK;
    };

    // ---- P cells: Cell_Geo_Cal with the closing rule and the corner patch
    const GridSpec GP{ni, nj, c->cls.xc, c->cls.yc, c->dx, c->dy, c->dx, c->dy, 0, 0};
    const GeoOut OP{F(SG_F_PHIS), F(SG_F_NXS), F(SG_F_NYS), F(SG_F_VOFS), F(SG_F_WLLH), F(SG_F_AE),
                    F(SG_F_AW),   F(SG_F_AN),  F(SG_F_AS),  W.pfc[0],    W.pfc[1],    W.pfc[2],
                    W.pfc[3]};
    hipLaunchKernelGGL(k_geo_cells, gc, b, 0, st, GP, S, OP);
    hipLaunchKernelGGL(k_geo_faces, gg, b, 0, st, GP, OP);
    SG_HIP(hipGetLastError());
    int rc = close_cells(OP, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_geo_wall, gc, b, 0, st, GP, OP);
    hipLaunchKernelGGL(k_geo_corner, gc, b, 0, st, GP, OP);
    SG_HIP(hipGetLastError());

    // ---- U and V cells: Cell_Geo_Cal without the closing rule
    // BGrid spacings of the interpolation routines: U east the P width of
    // column i+1, U north the V height; V east the U width, V north the P
    // height of row j+1
    double* dxn = W.next[0];
    double* dyn = W.next[1];
    if (!dxn) {
        W.next[0] = dxn = (double*)dev_alloc(c, ni * sizeof(double));
        W.next[1] = dyn = (double*)dev_alloc(c, nj * sizeof(double));
        if (!dxn || !dyn) {
            set_error("sg_geometry_rebuild: device allocation failed");
            return SG_ERR_ALLOC;
        }
        SG_HIP(hipMemcpyAsync(dxn, c->dx + 1, (ni - 1) * sizeof(double), hipMemcpyDeviceToDevice, st));
        SG_HIP(hipMemcpyAsync(dxn + ni - 1, c->dx + ni - 1, sizeof(double), hipMemcpyDeviceToDevice, st));
        SG_HIP(hipMemcpyAsync(dyn, c->dy + 1, (nj - 1) * sizeof(double), hipMemcpyDeviceToDevice, st));
        SG_HIP(hipMemcpyAsync(dyn + nj - 1, c->dy + nj - 1, sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    for (int w = 0; w < 2; ++w) {
        StagDev& D = c->stag[w];
        const GridSpec G{ni, nj, c->cls.xc, c->cls.yc, c->dx, c->dy, w == 0 ? c->udx : c->dx,
                         w == 0 ? c->dy : c->vdy, w == 0, w == 1};
        const GeoOut O{D.a[STG_PHIS], D.a[STG_NXS],  D.a[STG_NYS],  D.a[STG_VOFS], D.a[STG_WLLH],
                       D.a[STG_AE],   D.a[STG_AW],   D.a[STG_AN],   D.a[STG_AS],   D.a[STG_FCEX],
                       D.a[STG_FCEY], D.a[STG_FCNX], D.a[STG_FCNY]};
        hipLaunchKernelGGL(k_geo_cells, gc, b, 0, st, G, S, O);
        hipLaunchKernelGGL(k_geo_faces, gg, b, 0, st, G, O);
        SG_HIP(hipGetLastError());
        rc = close_cells(O, 0);
        if (rc) return rc;
        SG_HIP(hipMemcpyAsync(c->field[w == 0 ? SG_F_VOFSU : SG_F_VOFSV], D.a[STG_VOFS], ncell * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
    }

    // ---- momentum-exchange cells (closes P faces) and the U/V flags
    const MoexIn M{ni,      nj,      F(SG_F_VOFS), F(SG_F_VOFSU),           F(SG_F_VOFSV),
                   F(SG_F_AE), F(SG_F_AW), F(SG_F_AN), F(SG_F_AS),         W.moex[0],
                   W.moex[1], (int*)c->field[SG_F_UFLAG], (int*)c->field[SG_F_VFLAG]};
    hipLaunchKernelGGL(k_geo_moex, gc, b, 0, st, M);
    SG_HIP(hipGetLastError());

    // ---- NewCellFace and the interpolation factors
    for (int w = 0; w < 2; ++w) {
        StagDev& D = c->stag[w];
        StagIn s{};
        s.ni = ni;
        s.nj = nj;
        s.which = w;
        s.hx = w == 0 ? c->udx : c->dx;
        s.hy = w == 0 ? c->dy : c->vdy;
        s.sxn = w == 0 ? dxn : c->udx;
        s.syn = w == 0 ? c->vdy : dyn;
        s.pdx = c->dx;
        s.pdy = c->dy;
        s.pfx = w == 0 ? W.pfc[0] : W.pfc[2];
        s.pfy = w == 0 ? W.pfc[1] : W.pfc[3];
        s.moex = W.moex[w];
        for (int k = 0; k < STG_COUNT; ++k) s.a[k] = D.a[k];
        s.act = D.act;
        s.moex_out = D.moex;
        hipLaunchKernelGGL(k_stag_cells, gc, b, 0, st, s);
        hipLaunchKernelGGL(k_stag_east, gg, b, 0, st, s);
        hipLaunchKernelGGL(k_stag_north, gg, b, 0, st, s);
        SG_HIP(hipGetLastError());
    }
    SG_HIP(hipStreamSynchronize(st));
    c->built = 0;
    c->mg_ready = 0;
    return SG_OK;
}

}  // namespace sg
