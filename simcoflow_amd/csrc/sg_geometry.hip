// Cut-cell geometry rebuilt on the device for a solid given as a shape: the
// solid level set and normal on the P, U and V cells, the solid fractions
// (SolidVolumeFraction's role, Clsvof_Mod.f90:2553-2929, for a box,
// circle, ellipse or the oil boom's cylinder with its bar, Main.f90:173-182),
// the cell geometry of Cell_Geo_Cal / Edge_Geo_Cal
// (Cutcell_Mod.f90:34-337: edge solid fractions, wall length, apertures with
// the harmonic face averaging, solid cells closing their faces, cells closed
// on both sides of an axis becoming solid), DefineMomentumExchangeCell
// (:360-397) and the staggered-cell arrays the predictor reads (fluid
// centroids, open-face centres, interpolation factors in the simplified form
// the host builder uses for EastFaceInterpolationInf /
// NorthFaceInterpolationInf).  The result is the same, bit for bit, as the
// host builder (cases.build_geometry + cases.staggered_geometry) evaluated
// for the same shape: every kernel follows that builder's expressions in its
// order with FP contraction off.
//
// Two steps of the builder are sequential in loop order and are recast:
//  * the closing rule visits cells in (i, j) order, i outer; a cell's
//    decision depends only on the decisions of its west and south
//    neighbours (the cells visited before it that share a face with it), so
//    the decisions are the unique solution of a recurrence over an acyclic
//    dependency graph, found here by repeating a parallel pass until no
//    decision changes;
//  * DefineMomentumExchangeCell reads only the faces a cell itself writes,
//    so its cells are independent.
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

struct GeoOut {
    double *phi, *nx, *ny, *vofS, *wl;  // cell arrays
    double *aE, *aW, *aN, *aS;          // ghosted
};

struct RawAp {
    double *e, *w, *n, *s;  // ghosted, interior written, ghosts stay 0
};

__device__ inline size_t gix(int ni, int i, int j) { return (size_t)(i + 1) + (size_t)(j + 1) * (size_t)(ni + 2); }

// level set, solid fraction, edge solid fractions, raw wall length and
// apertures of every cell
__global__ void __launch_bounds__(256) k_geo_cells(GridSpec G, Shape S, GeoOut o, RawAp r) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= G.ni || j >= G.nj) return;
    const double X = G.shift_x ? G.xc[i] + 0.5 * G.dx[i] : G.xc[i];
    const double Y = G.shift_y ? G.yc[j] + 0.5 * G.dy[j] : G.yc[j];
    const double hx = G.hx[i], hy = G.hy[j];
    double phi, nx, ny;
    plane(S, X, Y, phi, nx, ny);
    double v = geo_frac(nx, ny, phi, hx, hy);
    if (v < G_VOFEPS) v = 0.0;
    if (v >= 1.0 - G_VOFEPS) v = 1.0;
    if (v > 1.0 - EPSI) v = 1.0;  // solid cells (Cutcell_Mod.f90:215-228)
    const double ex = 0.5 * hx, ey = 0.5 * hy;
    const double lsw = nx * (-ex) + ny * (-ey) + phi;
    const double lnw = nx * (-ex) + ny * ey + phi;
    const double lne = nx * ex + ny * ey + phi;
    const double lse = nx * ex + ny * (-ey) + phi;
    const double sW = edge_solid(lsw, lnw), sN = edge_solid(lnw, lne);
    const double sE = edge_solid(lne, lse), sS = edge_solid(lse, lsw);
    const double wx = (sE - sW) * hx, wy = (sN - sS) * hy;
    const size_t c = (size_t)i + (size_t)j * G.ni, g = gix(G.ni, i, j);
    o.phi[c] = phi;
    o.nx[c] = nx;
    o.ny[c] = ny;
    o.vofS[c] = v;
    o.wl[c] = sqrt(wx * wx + wy * wy);
    r.w[g] = 1.0 - sW;
    r.e[g] = 1.0 - sE;
    r.n[g] = 1.0 - sN;
    r.s[g] = 1.0 - sS;
}

__device__ inline double hmean(double a, double b) { return 2.0 * a * b / (a + b + G_ROUNDOFF); }

// apertures after the harmonic face averaging of i = 2..Isize-1,
// j = 2..Jsize-1 (Cutcell_Mod.f90:161-204); one thread per ghosted entry
__global__ void __launch_bounds__(256) k_geo_faces(int ni, int nj, RawAp r, GeoOut o) {
#pragma clang fp contract(off)
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
        return;
    }
    const bool avg = ni >= 3 && nj >= 3;
    const size_t up = g + (size_t)(ni + 2), dn = g - (size_t)(ni + 2);
    const bool row_in = j >= 1 && j <= nj - 2, col_in = i >= 1 && i <= ni - 2;
    o.aN[g] = (avg && col_in && row_in) ? hmean(r.n[g], r.s[up]) : r.n[g];
    o.aS[g] = (avg && col_in && j - 1 >= 1 && j - 1 <= nj - 2) ? hmean(r.n[dn], r.s[g]) : r.s[g];
    o.aE[g] = (avg && col_in && row_in) ? hmean(r.e[g], r.w[g + 1]) : r.e[g];
    o.aW[g] = (avg && i - 1 >= 1 && i - 1 <= ni - 2 && row_in) ? hmean(r.e[g - 1], r.w[g]) : r.w[g];
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

// solid cells close their faces and their neighbours' (writes are zeros only)
__global__ void __launch_bounds__(256) k_geo_solid(int ni, int nj, GeoOut o) {
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= ni || j >= nj) return;
    if (o.vofS[(size_t)i + (size_t)j * ni] > 1.0 - EPSI) close_faces(ni, gix(ni, i, j), o);
}

// one pass of the closing rule (Cutcell_Mod.f90:229-242): a cell is visited
// if it had a closed face when the rule started or its west or south
// neighbour closed; it closes if, with the faces those neighbours closed, it
// is shut on both sides along x or along y
__global__ void __launch_bounds__(256) k_geo_close(int ni, int nj, GeoOut o, const int* __restrict__ dold,
                                                   int* __restrict__ dnew, int* __restrict__ changed) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const bool fluid = o.vofS[c] < 1.0 - EPSI;
    int d = 0;
    if (fluid) {
        const double aW = o.aW[g], aE = o.aE[g], aN = o.aN[g], aS = o.aS[g];
        const bool cand = aW == 0.0 || aE == 0.0 || aN == 0.0 || aS == 0.0;
        const int dw = i > 0 ? dold[c - 1] : 0;
        const int ds = j > 0 ? dold[c - ni] : 0;
        if (cand || dw || ds) {
            const double w = dw ? 0.0 : aW;
            const double s = ds ? 0.0 : aS;
            d = (w + aE < G_ROUNDOFF || s + aN < G_ROUNDOFF) ? 1 : 0;
        }
    }
    dnew[c] = d;
    if (d != dold[c]) atomicOr(changed, 1);
}

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

struct MoexIn {
    int ni, nj;
    const double *vofS, *vofSU, *vofSV;
    double *aE, *aW, *aN, *aS;
    int *moexU, *moexV, *uflag, *vflag;
};

// DefineMomentumExchangeCell (Cutcell_Mod.f90:360-397) and the U/V flags
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
    m.uflag[c] = (su < 1.0 - EPSI && mu != 1) ? 1 : 0;
    m.vflag[c] = (sv < 1.0 - EPSI && mv != 1) ? 1 : 0;
}

// centroid of {nx x + ny y + phi >= 0} inside the hx*hy box centred at 0
// (polygon clipping), (0, 0) if fewer than three vertices remain
__device__ void fluid_centroid(double nx, double ny, double phi, double hx, double hy, double& cx, double& cy) {
#pragma clang fp contract(off)
    const double px[4] = {-hx / 2, hx / 2, hx / 2, -hx / 2};
    const double py[4] = {-hy / 2, -hy / 2, hy / 2, hy / 2};
    double ox[8], oy[8];
    int n = 0;
    for (int k = 0; k < 4; ++k) {
        const int q = (k + 1) & 3;
        const double fp = nx * px[k] + ny * py[k] + phi;
        const double fq = nx * px[q] + ny * py[q] + phi;
        if (fp >= 0) {
            ox[n] = px[k];
            oy[n] = py[k];
            ++n;
        }
        if ((fp < 0 && 0 <= fq) || (fq < 0 && 0 <= fp)) {
            const double t = fp / (fp - fq);
            ox[n] = px[k] + t * (px[q] - px[k]);
            oy[n] = py[k] + t * (py[q] - py[k]);
            ++n;
        }
    }
    cx = 0.0;
    cy = 0.0;
    if (n < 3) return;
    double a = 0.0, sx = 0.0, sy = 0.0;
    for (int k = 0; k < n; ++k) {
        const int q = k + 1 < n ? k + 1 : 0;
        const double cr = ox[k] * oy[q] - ox[q] * oy[k];
        a += cr;
        sx += (ox[k] + ox[q]) * cr;
        sy += (oy[k] + oy[q]) * cr;
    }
    if (fabs(a) < 1e-300) return;
    cx = sx / (3.0 * a);
    cy = sy / (3.0 * a);
}

struct StagIn {
    int ni, nj, which;   // 0: U cells, 1: V cells
    const double *hx, *hy;
    const double *sxn, *syn;  // centre distance to the next cell along x (per i) and y (per j)
    bool sxn_per_i;           // sxn indexed by i (else by j); syn likewise by j
    const int* moex;
    double* a[STG_COUNT];
    int* act;
    int* moex_out;
};

__device__ inline bool stag_cut(double v) { return v > EPSI && v < 1.0 - EPSI; }

// per cell: centroid, open-face centres, active rows, wall distance, and the
// default face factors
__global__ void __launch_bounds__(256) k_stag_cells(StagIn s) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const double hx = s.hx[i], hy = s.hy[j];
    const double v = s.a[STG_VOFS][c], nx = s.a[STG_NXS][c], ny = s.a[STG_NYS][c], phi = s.a[STG_PHIS][c];
    const bool cut = stag_cut(v);
    double cx = 0.0, cy = 0.0;
    if (cut) fluid_centroid(nx, ny, phi, hx, hy, cx, cy);
    s.a[STG_CCX][c] = cx;
    s.a[STG_CCY][c] = cy;
    s.a[STG_FCEX][g] = hx / 2;
    s.a[STG_FCNY][g] = hy / 2;
    s.a[STG_FCEY][g] = cut ? (0.5 - s.a[STG_AE][g] / 2.0) * np_sign(ny) * hy : 0.0;
    s.a[STG_FCNX][g] = cut ? (0.5 - s.a[STG_AN][g] / 2.0) * np_sign(nx) * hx : 0.0;
    int active = v < 1.0 - EPSI ? 1 : 0;
    if ((s.which == 0 && i == ni - 1) || (s.which == 1 && j == nj - 1)) active = 0;
    s.act[c] = active;
    const int mo = s.moex[c];
    s.moex_out[c] = mo;
    const double dl = fabs(cx * nx + cy * ny + phi) + 1.0e-24;
    s.a[STG_DELH][c] = mo == 1 ? 1.0e-20 : dl;
}

// per cell: its east and north faces (interpolation factors of cut faces)
__global__ void __launch_bounds__(256) k_stag_faces(StagIn s) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * TX + threadIdx.x;
    const int j = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (i >= ni || j >= nj) return;
    const size_t c = (size_t)i + (size_t)j * ni, g = gix(ni, i, j);
    const double hx = s.hx[i], hy = s.hy[j];
    const double sxn = s.sxn[i], syn = s.syn[j];
    const double* vofS = s.a[STG_VOFS];
    const double *ccx = s.a[STG_CCX], *ccy = s.a[STG_CCY];
    double SxE = sxn, EtaE = 0.5, AlE = 1.0, DAlE = 0.0;
    if (i + 1 < ni && s.act[c] && s.act[c + 1] && (stag_cut(vofS[c]) || stag_cut(vofS[c + 1]))) {
        const double sx = ccx[c + 1] + sxn - ccx[c];
        const double sy = ccy[c + 1] - ccy[c];
        SxE = sx;
        const double eta = fabs(s.a[STG_FCEX][g] - ccx[c]) / sx;
        EtaE = eta >= 1.0 ? 0.5 : eta;
        if (fabs(sy) > 1e-4 * hy) {
            const double al = 1.0 + fabs(sy) / (hy > 1e-300 ? hy : 1e-300);
            AlE = al < 2.0 ? al : 2.0;
            DAlE = sy * 0.5 * (s.a[STG_NYS][c] + s.a[STG_NYS][c + 1]) / (sx * hy);
        }
    }
    double SyN = syn, EtaN = 0.5, AlN = 1.0, DAlN = 0.0;
    if (j + 1 < nj && s.act[c] && s.act[c + ni] && (stag_cut(vofS[c]) || stag_cut(vofS[c + ni]))) {
        const double sy = ccy[c + ni] + syn - ccy[c];
        const double sx = ccx[c + ni] - ccx[c];
        SyN = sy;
        const double eta = fabs(s.a[STG_FCNY][g] - ccy[c]) / sy;
        EtaN = eta >= 1.0 ? 0.5 : eta;
        if (fabs(sx) >= 1e-4 * hx) {
            const double al = 1.0 + fabs(sx) / (hx > 1e-300 ? hx : 1e-300);
            AlN = al < 2.0 ? al : 2.0;
            DAlN = sx * 0.5 * (s.a[STG_NXS][c] + s.a[STG_NXS][c + ni]) / (sy * hx);
        }
    }
    s.a[STG_SXE][g] = SxE;
    s.a[STG_ETAE][g] = EtaE;
    s.a[STG_ALE][g] = AlE;
    s.a[STG_DALE][g] = DAlE;
    s.a[STG_SYN][g] = SyN;
    s.a[STG_ETAN][g] = EtaN;
    s.a[STG_ALN][g] = AlN;
    s.a[STG_DALN][g] = DAlN;
}

// ghost entries of the face-factor and face-centre arrays take the value of
// the nearest interior entry (rows copied first, then columns)
__global__ void __launch_bounds__(256) k_stag_ghosts(StagIn s) {
    const int gi = blockIdx.x * TX + threadIdx.x;
    const int gj = blockIdx.y * TY + threadIdx.y;
    const int ni = s.ni, nj = s.nj;
    if (gi >= ni + 2 || gj >= nj + 2) return;
    if (gi >= 1 && gi <= ni && gj >= 1 && gj <= nj) return;
    const int si = gi < 1 ? 1 : (gi > ni ? ni : gi);
    const int sj = gj < 1 ? 1 : (gj > nj ? nj : gj);
    const size_t g = (size_t)gi + (size_t)gj * (size_t)(ni + 2), src = (size_t)si + (size_t)sj * (size_t)(ni + 2);
    const int ks[] = {STG_FCEX, STG_FCEY, STG_FCNX, STG_FCNY, STG_SXE, STG_ETAE, STG_ALE, STG_DALE,
                      STG_SYN,  STG_ETAN, STG_ALN,  STG_DALN};
    for (int k : ks) s.a[k][g] = s.a[k][src];
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
    if (!W.raw[0]) {
        for (double*& p : W.raw) p = (double*)dev_alloc(c, nghost * sizeof(double));
        for (int*& p : W.close) p = (int*)dev_alloc(c, ncell * sizeof(int));
        W.changed = (int*)dev_alloc(c, sizeof(int));
        W.moex[0] = (int*)dev_alloc(c, ncell * sizeof(int));
        W.moex[1] = (int*)dev_alloc(c, ncell * sizeof(int));
        if (!W.raw[3] || !W.close[1] || !W.changed || !W.moex[0] || !W.moex[1]) {
            set_error("sg_geometry_rebuild: device allocation failed");
            return SG_ERR_ALLOC;
        }
    }
    hipStream_t st = c->stream;
    const Shape S{shape->kind, shape->cx, shape->cy, shape->a, shape->b, shape->c};
    const RawAp R{W.raw[0], W.raw[1], W.raw[2], W.raw[3]};
    const dim3 gc = grid2d(ni, nj), gg = grid2d(ni + 2, nj + 2), b = block2d();
    auto F = [&](int f) { return (double*)c->field[f]; };

    // ---- P cells (Grid_Preprocess on PCell)
    const GridSpec GP{ni, nj, c->cls.xc, c->cls.yc, c->dx, c->dy, c->dx, c->dy, 0, 0};
    const GeoOut OP{F(SG_F_PHIS), F(SG_F_NXS), F(SG_F_NYS), F(SG_F_VOFS), F(SG_F_WLLH),
                    F(SG_F_AE),   F(SG_F_AW),  F(SG_F_AN),  F(SG_F_AS)};
    hipLaunchKernelGGL(k_geo_cells, gc, b, 0, st, GP, S, OP, R);
    hipLaunchKernelGGL(k_geo_faces, gg, b, 0, st, ni, nj, R, OP);
    hipLaunchKernelGGL(k_geo_solid, gc, b, 0, st, ni, nj, OP);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemsetAsync(W.close[0], 0, ncell * sizeof(int), st));
    int cur = 0, passes = 0;
    for (;;) {
        int changed = 0;
        SG_HIP(hipMemsetAsync(W.changed, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_geo_close, gc, b, 0, st, ni, nj, OP, (const int*)W.close[cur], W.close[1 - cur],
                           W.changed);
        SG_HIP(hipGetLastError());
        SG_HIP(hipMemcpyAsync(&changed, W.changed, sizeof(int), hipMemcpyDeviceToHost, st));
        SG_HIP(hipStreamSynchronize(st));
        cur = 1 - cur;
        if (!changed) break;
        if (++passes > MAX_CLOSE_PASSES) {
            set_error("sg_geometry_rebuild: the closing rule did not settle");
            return SG_ERR_STATE;
        }
    }
    hipLaunchKernelGGL(k_geo_close_apply, gc, b, 0, st, ni, nj, OP, (const int*)W.close[cur]);
    hipLaunchKernelGGL(k_geo_wall, gc, b, 0, st, GP, OP);
    SG_HIP(hipGetLastError());

    // ---- U and V cells: solid fractions, apertures (no closing rule)
    // spacing to the next cell: U along x the P width of column i+1, along
    // y dy; V along x the U width, along y the P height of row j+1
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
        const GeoOut O{D.a[STG_PHIS], D.a[STG_NXS], D.a[STG_NYS], D.a[STG_VOFS], D.a[STG_WLLH],
                       D.a[STG_AE],   D.a[STG_AW],  D.a[STG_AN],  D.a[STG_AS]};
        hipLaunchKernelGGL(k_geo_cells, gc, b, 0, st, G, S, O, R);
        hipLaunchKernelGGL(k_geo_faces, gg, b, 0, st, ni, nj, R, O);
        hipLaunchKernelGGL(k_geo_solid, gc, b, 0, st, ni, nj, O);
        hipLaunchKernelGGL(k_geo_wall, gc, b, 0, st, G, O);
        SG_HIP(hipGetLastError());
        SG_HIP(hipMemcpyAsync(c->field[w == 0 ? SG_F_VOFSU : SG_F_VOFSV], D.a[STG_VOFS], ncell * sizeof(double),
                              hipMemcpyDeviceToDevice, st));
    }

    // ---- momentum-exchange cells (closes P faces) and the U/V flags
    const MoexIn M{ni,      nj,      F(SG_F_VOFS), F(SG_F_VOFSU),           F(SG_F_VOFSV),
                   F(SG_F_AE), F(SG_F_AW), F(SG_F_AN), F(SG_F_AS),         W.moex[0],
                   W.moex[1], (int*)c->field[SG_F_UFLAG], (int*)c->field[SG_F_VFLAG]};
    hipLaunchKernelGGL(k_geo_moex, gc, b, 0, st, M);
    SG_HIP(hipGetLastError());

    // ---- staggered-cell arrays of the predictor
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
        s.moex = W.moex[w];
        for (int k = 0; k < STG_COUNT; ++k) s.a[k] = D.a[k];
        s.act = D.act;
        s.moex_out = D.moex;
        hipLaunchKernelGGL(k_stag_cells, gc, b, 0, st, s);
        hipLaunchKernelGGL(k_stag_faces, gc, b, 0, st, s);
        hipLaunchKernelGGL(k_stag_ghosts, gg, b, 0, st, s);
        SG_HIP(hipGetLastError());
    }
    SG_HIP(hipStreamSynchronize(st));
    c->built = 0;
    c->mg_ready = 0;
    return SG_OK;
}

}  // namespace sg
