// Momentum predictor on gfx950 (PredictorUV_mod.f90, Predictor_UV :28-497).
//
//   k_conv        per U / V cell: convective fluxes of its faces (MUSCL +
//                 kinetic flux, or the first-step upwind form), face
//                 densities from the half-step PLIC volume and the
//                 density-weighted convective fluxes; then the
//                 density-corrected convective divergence, Adams-Bashforth
//                 extrapolation with Flux_n1, closed-cell zeroing, the
//                 u^n dx dy / dt term
//   k_mix_visc    viscous conductances and cross-diffusion terms of its
//                 faces at time n; small-cell mixing (Mixing_Procedure
//                 :2127-2197) gathered in the reference's scatter order,
//                 explicit half of the viscous term, moving-wall friction
//   k_rows        viscous conductances with the vof after the transport;
//                 the implicit momentum rows, right-hand sides and Dp
//                 (SetMatrix :622-768, SetVectors :770-845)
//   BiCGStab      the two momentum systems (sg_momentum.hip)
//   k_scatter, k_finish  Pred%u/v, ghost copies of Dp, open-top v
//
// Each of the three cell kernels computes the face values its tile of U/V
// cells reads into LDS: no face flux goes through memory.  Face (i,j) of
// direction EW is the west face of U/V cell (i,j), i = 1..ni+1; NS face
// (i,j) the south face, j = 1..nj+1.  Compiled with FP contraction off:
// every operation rounds as the reference writes it.
#include "sg_internal.hpp"
#include "sg_poly.hpp"

namespace sg {

namespace {

constexpr double BETAVIS = 0.5;  // StateVariables.f90:28

struct StagV {  // one staggered grid on the device
    const double *vofS, *nxS, *nyS, *phiS, *WlLh, *delh, *ccx, *ccy;
    const int *moex, *act;
    const double *aE, *aW, *aN, *aS, *EtaE, *EtaN, *AlE, *AlN, *DAlE, *DAlN, *SxE, *SyN;
    const double *fcex, *fcey, *fcnx, *fcny;
    const double *vofo, *vofn;  // U/VCellO%vof, U/VCell%vof
};

struct PredG {
    int ni, nj;
    const double *dx, *dy, *udx, *vdy;
    StagV U, V;
    const double *pvof, *pphi, *pnx, *pny, *pvofS, *pnxS, *pnyS, *pphiS;  // P cells at time n
    const double *u, *v;
    double rl, rg, rr, nul, nug, nur, Re, Fr, dt, vb;
    int first;
};

__device__ __forceinline__ size_t ci(const PredG& G, int i, int j) {
    return (size_t)(i - 1) + (size_t)(j - 1) * (size_t)G.ni;
}
__device__ __forceinline__ size_t gi(const PredG& G, int i, int j) {
    return (size_t)i + (size_t)j * (size_t)(G.ni + 2);
}
__device__ __forceinline__ double sq(double x) { return x * x; }
__device__ __forceinline__ double pow5(double x) { return ((x * x) * (x * x)) * x; }
__device__ __forceinline__ double fsign(double a, double b) { return copysign(fabs(a), b); }

#define UU(i, j) G.u[gi(G, i, j)]
#define VV(i, j) G.v[gi(G, i, j)]
#define PDX(i) G.dx[(i)-1]
#define PDY(j) G.dy[(j)-1]
#define UDX(i) G.udx[(i)-1]
#define UDY(j) G.dy[(j)-1]
#define VDX(i) G.dx[(i)-1]
#define VDY(j) G.vdy[(j)-1]

// superbee MUSCL limiter (MUSCLLimiter :2354-2367, opt 2)
__device__ __forceinline__ double muscl(double x, double tolim) {
    return fmax(fmax(0.0, fmin(tolim * x, 1.0)), fmin(x, tolim));
}

// PLIC liquid fraction (frac :2215-2257)
__device__ double pfrac(double nx, double ny, double d, double dx, double dy) {
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

// densities of FaceDensityFlux / DensityBasedConvectiveFlux in their two
// operand orders
__device__ __forceinline__ double mix_rho(const PredG& G, double vf, double vs) {
    return ((1.0 - vf - vs) * G.rg / G.rr + vf * G.rl / G.rr);
}
__device__ __forceinline__ double rho_mix(const PredG& G, double vf, double vs) {
    return (vf * G.rl / G.rr + (1.0 - vf - vs) * G.rg / G.rr);
}
__device__ __forceinline__ double nu_face(const PredG& G, double vf) {
    return ((1.0 - vf) * G.nug / G.nur + vf * G.nul / G.nur);
}
__device__ __forceinline__ double edge_mix(const PredG& G, double vf, double vs) {
    const double tol = 1.0e-24;
    return ((1.0 - vf / (1.0 - vs + tol)) * G.nug / G.nur + vf / (1.0 - vs + tol) * G.nul / G.nur);
}
// vofS > EPSI in any of four cells; the four loads issue together
__device__ __forceinline__ bool any_solid4(const double* vofS, size_t a, size_t b, size_t c, size_t d) {
    const double s0 = vofS[a], s1 = vofS[b], s2 = vofS[c], s3 = vofS[d];
    return (s0 > EPSI) | (s1 > EPSI) | (s2 > EPSI) | (s3 > EPSI);
}
__device__ __forceinline__ void upwind(double w, double& p, double& m) {
    p = 0.5 * (w + fabs(w));
    m = 0.5 * (w - fabs(w));
}

// ---- convective fluxes of one face --------------------------------------------

// EW face i of row j: both components (HighOrderConvectiveFluxForXDir
// :1186-1305 or ModifiedConvectiveFlux :888-997 at the first step)
__device__ void cflux_ew(const PredG& G, int i, int j, double& f1, double& f2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const int ni = G.ni;
    if (G.first) {
        double uw = 0.5 * (UU(i, j) + UU(i - 1, j));
        if (i == 1) {
            f1 = sq(uw) * UDY(j) * Uc.aW[gi(G, i, j)];
        } else if (i >= ni) {
            f1 = sq(uw) * UDY(j) * Uc.aE[gi(G, i - 1, j)];
        } else {
            const double eta = Uc.EtaE[gi(G, i - 1, j)];
            uw = (1.0 - eta) * UU(i - 1, j) + eta * UU(i, j);
            double p, m;
            upwind(uw, p, m);
            f1 = (p * UU(i - 1, j) + m * UU(i, j)) * Uc.AlE[gi(G, i - 1, j)] * UDY(j) * Uc.aW[gi(G, i, j)];
        }
        uw = 0.5 * (UU(i - 1, j + 1) + UU(i - 1, j));
        if (i > ni) {
            f2 = uw * 0.5 * (VV(i - 1, j) + VV(i, j)) * VDY(j) * Vc.aE[gi(G, i - 1, j)];
        } else if (i == 1) {
            f2 = uw * 0.5 * (VV(i - 1, j) + VV(i, j)) * VDY(j) * Vc.aW[gi(G, i, j)];
        } else {
            const double Sy = Uc.SyN[gi(G, i - 1, j)];
            double eta = fabs(Vc.fcey[gi(G, i - 1, j)] + PDY(j) / 2.0 - Uc.ccy[ci(G, i - 1, j)]) / Sy;
            if (fabs(eta) >= 1.0) eta = 0.5;
            uw = (1.0 - eta) * UU(i - 1, j) + eta * UU(i - 1, j + 1);
            double p, m;
            upwind(uw, p, m);
            f2 = (p * VV(i - 1, j) + m * VV(i, j)) * Vc.aW[gi(G, i, j)] * VDY(j);
        }
        return;
    }
    if (i == 1) {
        const double ul = UU(0, j), ur = UU(1, j);
        const double alr = fmax(fabs(ur), fabs(ul));
        f1 = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc.aW[gi(G, 1, j)] * UDY(j);
        const double vl = VV(0, j), vr = VV(1, j);
        const double a2 = 0.5 * (UU(0, j) + UU(0, j + 1));
        f2 = 0.5 * (a2 * vr + a2 * vl - fabs(a2) * (vr - vl)) * Vc.aW[gi(G, 1, j)] * VDY(j);
        return;
    }
    if (i == ni + 1) {
        const double ul = UU(ni, j), ur = UU(ni + 1, j);
        const double alr = fmax(fabs(ur), fabs(ul));
        f1 = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc.aE[gi(G, ni, j)] * UDY(j);
        const double vl = VV(ni, j), vr = VV(ni + 1, j);
        const double a2 = 0.5 * (UU(ni, j) + UU(ni, j + 1));
        f2 = 0.5 * (a2 * vr + a2 * vl - fabs(a2) * (vr - vl)) * Vc.aE[gi(G, ni, j)] * VDY(j);
        return;
    }
    const double tol = 1.0e-24;
    const int ip = i + 1 < ni ? i + 1 : ni;
    double tolim = fmin(2.0 * UDX(i - 1) / PDX(i), 2.0 * UDX(i) / PDX(i));
    double ri = ((UU(i - 1, j) - UU(i - 2, j)) / PDX(i - 1)) / ((UU(i, j) - UU(i - 1, j)) / PDX(i));
    double omei = muscl(ri, tolim) * (UU(i, j) - UU(i - 1, j)) / PDX(i);
    double ri1 = ((UU(i, j) - UU(i - 1, j)) / PDX(i)) / ((UU(i + 1, j) - UU(i, j)) / PDX(ip));
    double omei1 = muscl(ri1, tolim) * (UU(i + 1, j) - UU(i, j)) / PDX(ip);
    double ul = UU(i - 1, j) + 0.5 * (PDX(i)) * omei;
    double ur = UU(i, j) - 0.5 * (PDX(i)) * omei1;
    double alr = fmax(fabs(ur), fabs(ul));
    f1 = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc.aW[gi(G, i, j)] * UDY(j);
    if (i > 2 && i < ni) {
        if (any_solid4(Uc.vofS, ci(G, i - 2, j), ci(G, i - 1, j), ci(G, i, j), ci(G, i + 1, j))) {
            const double eta = Uc.EtaE[gi(G, i - 1, j)];
            const double uw = (1 - eta) * UU(i - 1, j) + eta * UU(i, j);
            double p, m;
            upwind(uw, p, m);
            f1 = (p * UU(i - 1, j) + m * UU(i, j)) * Uc.AlE[gi(G, i - 1, j)] * UDY(j) * Uc.aW[gi(G, i, j)];
        }
    }
    const int im2 = i - 2 > 1 ? i - 2 : 1;
    tolim = fmin(2.0 * VDX(i - 1) / UDX(i - 1), 2.0 * VDX(i) / UDX(i - 1));
    ri = ((VV(i - 1, j) - VV(i - 2, j)) / UDX(im2)) / ((VV(i, j) - VV(i - 1, j)) / UDX(i - 1));
    omei = muscl(ri, tolim) * (VV(i, j) - VV(i - 1, j)) / UDX(i - 1);
    ri1 = ((VV(i, j) - VV(i - 1, j)) / UDX(i - 1)) / ((VV(i + 1, j) - VV(i, j)) / UDX(i));
    omei1 = muscl(ri1, tolim) * (VV(i + 1, j) - VV(i, j)) / UDX(i);
    const double vl = VV(i - 1, j) + 0.5 * (PDX(i - 1)) * omei;
    const double vr = VV(i, j) - 0.50 * (PDX(i)) * omei1;
    alr = 0.5 * (UU(i - 1, j) + UU(i - 1, j + 1));
    f2 = 0.5 * (alr * vr + alr * vl - fabs(alr) * (vr - vl)) * Vc.aW[gi(G, i, j)] * VDY(j);
    if (i > 2 && i < ni) {
        if (any_solid4(Vc.vofS, ci(G, i - 2, j), ci(G, i - 1, j), ci(G, i, j), ci(G, i + 1, j))) {
            if (Vc.aW[gi(G, i, j)] < 0.5) {
                const size_t c1 = ci(G, i - 1, j);
                const double delhec =
                    fabs(Vc.fcex[gi(G, i - 1, j)] * Vc.nxS[c1] + Vc.fcey[gi(G, i - 1, j)] * Vc.nyS[c1] + Vc.phiS[c1]);
                double delh, uw;
                const size_t a = ci(G, i - 1, j + 1), b = ci(G, i - 1, j);
                if (Uc.moex[a] != 1 && Uc.vofS[a] < 1.0 - EPSI) {
                    delh = fabs(Uc.ccx[a] * Uc.nxS[a] + Uc.ccy[a] * Uc.nyS[a] + Uc.phiS[a]) + tol;
                    uw = UU(i - 1, j + 1);
                } else if (Uc.moex[b] != 1 && Uc.vofS[b] < 1.0 - EPSI) {
                    delh = fabs(Uc.ccx[b] * Uc.nxS[b] + Uc.ccy[b] * Uc.nyS[b] + Uc.phiS[b]) + tol;
                    uw = UU(i - 1, j);
                } else {
                    delh = delhec + tol;
                    uw = 0.0;
                }
                double p, m;
                upwind(uw, p, m);
                f2 = (p * VV(i - 1, j) + m * VV(i, j)) * delhec / delh * Vc.aW[gi(G, i, j)] * VDY(j);
            } else {
                const double Sy = Uc.SyN[gi(G, i - 1, j)];
                double eta = fabs(Vc.fcey[gi(G, i - 1, j)] + VDY(j) / 2.0 - Uc.ccy[ci(G, i - 1, j)]) / Sy;
                if (fabs(eta) >= 1.0) eta = 0.5;
                const double uw = (1.0 - eta) * UU(i - 1, j) + eta * UU(i - 1, j + 1);
                double p, m;
                upwind(uw, p, m);
                f2 = (p * VV(i - 1, j) + m * VV(i, j)) * Vc.aW[gi(G, i, j)] * VDY(j);
            }
        }
    }
}

// NS face j of column i (HighOrderConvectiveFluxForYDir :1307-1429, or
// ModifiedConvectiveFlux at the first step)
__device__ void cflux_ns(const PredG& G, int i, int j, double& f1, double& f2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const int nj = G.nj;
    const double vb = G.vb;
    if (G.first) {
        double vs = 0.5 * (VV(i, j - 1) + VV(i + 1, j - 1));
        if (j > nj) {
            f1 = vs * 0.5 * (UU(i, j) + UU(i, j - 1)) * UDX(i) * Uc.aN[gi(G, i, j - 1)];
        } else if (j == 1) {
            f1 = vs * 0.5 * (UU(i, j) + UU(i, j - 1)) * UDX(i) * Uc.aS[gi(G, i, j)];
        } else {
            const double Sx = Vc.SxE[gi(G, i, j - 1)];
            double eta = fabs(Uc.fcnx[gi(G, i, j - 1)] + PDX(i) / 2.0 - Vc.ccx[ci(G, i, j - 1)]) / Sx;
            if (fabs(eta) >= 1.0) eta = 0.5;
            vs = (1.0 - eta) * VV(i, j - 1) + eta * VV(i + 1, j - 1);
            double p, m;
            upwind(vs, p, m);
            f1 = (p * UU(i, j - 1) + m * UU(i, j)) * Uc.aS[gi(G, i, j)] * UDX(i);
        }
        vs = 0.5 * (VV(i, j) + VV(i, j - 1));
        if (j >= nj) {
            f2 = sq(vs) * VDX(i) * Vc.aN[gi(G, i, j - 1)];
        } else if (j == 1) {
            f2 = sq(vs) * VDX(i) * Vc.aS[gi(G, i, j)];
        } else {
            const double eta = Vc.EtaN[gi(G, i, j - 1)];
            vs = (1.0 - eta) * VV(i, j - 1) + eta * VV(i, j);
            vs = ((vs - vb) * Vc.AlN[gi(G, i, j - 1)] + vb);
            double p, m;
            upwind(vs, p, m);
            f2 = (p * VV(i, j - 1) + m * VV(i, j)) * Vc.aS[gi(G, i, j)] * VDX(i);
        }
        return;
    }
    if (j == 1) {
        const double ul = UU(i, 0), ur = UU(i, 1);
        const double alr = 0.5 * (VV(i, 0) + VV(i + 1, 0));
        // `alr*ur*+alr*ul` (:1322), as gfortran parses it
        f1 = 0.5 * (((alr * ur) * alr) * ul - fabs(alr) * (ur - ul)) * Uc.aS[gi(G, i, 1)] * UDX(i);
        const double vl = VV(i, 0), vr = VV(i, 1);
        const double a2 = fmax(fabs(vl), fabs(vr));
        f2 = 0.5 * (sq(vr) + sq(vl) - fabs(a2) * (vr - vl)) * Vc.aS[gi(G, i, 1)] * VDX(i);
        return;
    }
    if (j == nj + 1) {
        const double ul = UU(i, nj), ur = UU(i, nj + 1);
        const double alr = 0.5 * (VV(i, nj) + VV(i + 1, nj));
        f1 = 0.5 * (alr * ur + alr * ul - fabs(alr) * (ur - ul)) * Uc.aN[gi(G, i, nj)] * UDX(i);
        const double vl = VV(i, nj), vr = VV(i, nj + 1);
        const double a2 = fmax(fabs(vl), fabs(vr));
        f2 = 0.5 * (sq(vr) + sq(vl) - fabs(a2) * (vr - vl)) * Vc.aN[gi(G, i, nj)] * VDX(i);
        return;
    }
    const double tol = 1.0e-24;
    const int jm2 = j - 2 > 1 ? j - 2 : 1;
    const int jp = j + 1 < nj ? j + 1 : nj;
    double tolim = fmin(2.0 * UDY(j - 1) / VDY(j - 1), 2.0 * UDY(j) / VDY(j - 1));
    double ri = ((UU(i, j - 1) - UU(i, j - 2)) / VDY(jm2)) / ((UU(i, j) - UU(i, j - 1)) / VDY(j - 1));
    double omei = muscl(ri, tolim) * (UU(i, j) - UU(i, j - 1)) / VDY(j - 1);
    double ri1 = ((UU(i, j) - UU(i, j - 1)) / VDY(j - 1)) / ((UU(i, j + 1) - UU(i, j)) / VDY(j));
    double omei1 = muscl(ri1, tolim) * (UU(i, j + 1) - UU(i, j)) / VDY(j);
    const double ul = UU(i, j - 1) + 0.5 * (PDY(j - 1)) * omei;
    const double ur = UU(i, j) - 0.5 * (PDY(j)) * omei1;
    double alr = 0.5 * (VV(i, j - 1) + VV(i + 1, j - 1));
    f1 = 0.5 * (alr * ur + alr * ul - fabs(alr) * (ur - ul)) * Uc.aS[gi(G, i, j)] * UDX(i);
    if (j > 2 && j < nj) {
        if (any_solid4(Uc.vofS, ci(G, i, j - 2), ci(G, i, j - 1), ci(G, i, j), ci(G, i, j + 1))) {
            if (Uc.aS[gi(G, i, j)] < 0.5) {
                const size_t c1 = ci(G, i, j - 1);
                const double delhec =
                    fabs(Uc.fcnx[gi(G, i, j - 1)] * Uc.nxS[c1] + Uc.fcny[gi(G, i, j - 1)] * Uc.nyS[c1] + Uc.phiS[c1]);
                double delh, vs;
                const size_t a = ci(G, i, j - 1), b = ci(G, i + 1, j - 1);
                if (Vc.moex[a] != 1 && Vc.vofS[a] < 1.0 - EPSI) {
                    delh = fabs(Vc.ccx[a] * Vc.nxS[a] + Vc.ccy[a] * Vc.nyS[a] + Vc.phiS[a]) + tol;
                    vs = VV(i, j - 1);
                } else if (Vc.moex[b] != 1 && Vc.vofS[b] < 1.0 - EPSI) {
                    delh = fabs(Vc.ccx[b] * Vc.nxS[b] + Vc.ccy[b] * Vc.nyS[b] + Vc.phiS[b]) + tol;
                    vs = VV(i + 1, j - 1);
                } else {
                    delh = delhec + tol;
                    vs = 0.0;
                }
                vs = vb + (vs - vb) * delhec / delh;
                double p, m;
                upwind(vs, p, m);
                f1 = (p * UU(i, j - 1) + m * UU(i, j)) * delhec / delh * Uc.aS[gi(G, i, j)] * UDX(i);
            } else {
                const double Sx = Vc.SxE[gi(G, i, j - 1)];
                double eta = fabs(Uc.fcnx[gi(G, i, j - 1)] + 0.5 * UDY(j - 1) - Vc.ccx[ci(G, i, j - 1)]) / Sx;
                if (fabs(eta) >= 1.0) eta = 0.5;
                const double vs = (1.0 - eta) * VV(i, j - 1) + eta * VV(i + 1, j - 1);
                double p, m;
                upwind(vs, p, m);
                f1 = (p * UU(i, j - 1) + m * UU(i, j)) * Uc.aS[gi(G, i, j)] * UDX(i);
            }
        }
    }
    tolim = fmin(2.0 * VDY(j - 1) / PDY(j), 2.0 * VDY(j) / PDY(j));
    ri = ((VV(i, j - 1) - VV(i, j - 2)) / PDY(j - 1)) / ((VV(i, j) - VV(i, j - 1)) / PDY(j));
    omei = muscl(ri, tolim) * (VV(i, j) - VV(i, j - 1)) / PDY(j - 1);
    ri1 = ((VV(i, j) - VV(i, j - 1)) / PDY(j)) / ((VV(i, j + 1) - VV(i, j)) / PDY(jp));
    omei1 = muscl(ri1, tolim) * (VV(i, j + 1) - VV(i, j)) / PDY(jp);
    const double vl = VV(i, j - 1) + 0.5 * (PDY(j)) * omei;
    const double vr = VV(i, j) - 0.5 * (PDY(j)) * omei1;
    alr = fmax(fabs(vl), fabs(vr));
    f2 = 0.5 * (vr * vr + vl * vl - fabs(alr) * (vr - vl)) * Vc.aS[gi(G, i, j)] * VDX(i);
    if (j > 2 && j < nj) {
        if (any_solid4(Vc.vofS, ci(G, i, j - 2), ci(G, i, j - 1), ci(G, i, j), ci(G, i, j + 1))) {
            const double eta = Vc.EtaN[gi(G, i, j - 1)];
            double vs = (1.0 - eta) * VV(i, j - 1) + eta * VV(i, j);
            vs = (vb + (vs - vb) * Vc.AlN[gi(G, i, j - 1)]);
            double p, m;
            upwind(vs, p, m);
            f2 = (p * VV(i, j - 1) + m * VV(i, j)) * Vc.aS[gi(G, i, j)] * VDX(i);
        }
    }
}

// ---- densities and density-weighted fluxes of one face ----------------------

// EW face: MaFlux (m1,m2) and ConFlux (k1,k2) (FaceDensityFlux :1884-1949,
// DensityBasedConvectiveFlux :1762-1812)
__device__ void dflux_ew(const PredG& G, int i, int j, double& m1, double& m2, double& k1, double& k2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const int ni = G.ni;
    const double dt = G.dt;
    // U component
    double uw = 0.5 * (UU(i, j) + UU(i - 1, j));
    double w1 = 0.0;
    if (i == 1) {
        if (Uc.aW[gi(G, 1, j)] >= EPSI) w1 = mix_rho(G, G.pvof[ci(G, i, j)], G.pvofS[ci(G, i, j)]);
        m1 = uw * w1 * UDY(j);
    } else if (i >= ni) {
        if (Uc.aE[gi(G, ni, j)] >= EPSI) w1 = mix_rho(G, G.pvof[ci(G, ni, j)], G.pvofS[ci(G, ni, j)]);
        m1 = uw * w1 * UDY(j);
    } else {
        // operands of every branch below, loaded before the first test
        const size_t c = ci(G, i, j);
        const double aw = Uc.aW[gi(G, i, j)], eta = Uc.EtaE[gi(G, i - 1, j)], al = Uc.AlE[gi(G, i - 1, j)];
        const double pv = G.pvof[c], pvs = G.pvofS[c];
        if (aw >= EPSI) {
            uw = (1.0 - eta) * UU(i - 1, j) + eta * UU(i, j);
            if (pv > 1.0 - EPSI || pv < EPSI) {
                w1 = mix_rho(G, pv, pvs);
            } else if (pvs < EPSI) {
                const double volf = pfrac(G.pnx[c], G.pny[c], G.pphi[c] + uw * dt / 2.0 * G.pnx[c], fabs(uw * dt), PDY(j));
                w1 = (volf * G.rl / G.rr + (1.0 - volf - 0.0) * G.rg / G.rr);
            } else {
                double vols, volf;
                cut_cell_fractions(G.pnxS[c], G.pnyS[c], G.pphiS[c] + uw * dt / 2.0 * G.pnxS[c], G.pnx[c],
                                   G.pny[c], G.pphi[c] + uw * dt / 2.0 * G.pnx[c], fabs(uw * dt), PDY(j), vols,
                                   volf);
                w1 = (volf * G.rl / G.rr + (1.0 - volf - vols) * G.rg / G.rr);
            }
        }
        m1 = uw * al * w1 * UDY(j);
    }
    // density-weighted U momentum flux
    {
        double w = 0.5 * (UU(i, j) + UU(i - 1, j));
        if (i == 1 || i >= ni) {
            k1 = sq(w) * w1 * UDY(j);
        } else {
            const double eta = Uc.EtaE[gi(G, i - 1, j)];
            w = (1.0 - eta) * UU(i - 1, j) + eta * UU(i, j);
            double p, m;
            upwind(w, p, m);
            k1 = (p * UU(i - 1, j) + m * UU(i, j)) * Uc.AlE[gi(G, i - 1, j)] * w1 * UDY(j);
        }
    }
    // V component (convective velocity u, scalar v)
    uw = 0.5 * (UU(i - 1, j + 1) + UU(i - 1, j));
    m2 = 0.0;
    k2 = 0.0;
    if (i > ni) {
        if (Vc.aE[gi(G, ni, j)] >= EPSI) {
            const double r = rho_mix(G, Vc.vofo[ci(G, ni, j)], Vc.vofS[ci(G, ni, j)]);
            m2 = uw * r * VDY(j);
            k2 = uw * 0.5 * (VV(i - 1, j) + VV(i, j)) * r * VDY(j);
        }
    } else if (i == 1) {
        if (Vc.aW[gi(G, 1, j)] >= EPSI) {
            const double r = rho_mix(G, Vc.vofo[ci(G, i, j)], Vc.vofS[ci(G, i, j)]);
            m2 = uw * r * VDY(j);
            k2 = uw * 0.5 * (VV(i - 1, j) + VV(i, j)) * r * VDY(j);
        }
    } else {
        const double Sy = G.U.SyN[gi(G, i - 1, j)];
        double eta = fabs(Vc.fcey[gi(G, i - 1, j)] + PDY(j) / 2.0 - G.U.ccy[ci(G, i - 1, j)]) / Sy;
        if (fabs(eta) >= 1.0) eta = 0.5;
        const double w = (1.0 - eta) * UU(i - 1, j) + eta * UU(i - 1, j + 1);
        double p, m;
        upwind(w, p, m);
        const size_t ca = ci(G, i - 1, j), cb = ci(G, i, j);
        const double fa = Vc.vofo[ca], sa = Vc.vofS[ca], fb = Vc.vofo[cb], sb = Vc.vofS[cb];
        if (Vc.aW[gi(G, i, j)] >= EPSI) {
            const double ra = rho_mix(G, fa, sa);
            const double rb = rho_mix(G, fb, sb);
            m2 = (p * ra + m * rb) * VDY(j);
            k2 = (p * VV(i - 1, j) * ra + m * VV(i, j) * rb) * VDY(j);
        }
    }
}

// NS face (FaceDensityFlux :1951-2027, DensityBasedConvectiveFlux :1816-1867)
__device__ void dflux_ns(const PredG& G, int i, int j, double& m1, double& m2, double& k1, double& k2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const int nj = G.nj;
    const double dt = G.dt, vb = G.vb;
    // U component (convective velocity v, scalar u)
    double vs = 0.5 * (VV(i, j - 1) + VV(i + 1, j - 1));
    m1 = 0.0;
    k1 = 0.0;
    if (j > nj) {
        if (Uc.aN[gi(G, i, nj)] >= EPSI) {
            const double r = rho_mix(G, Uc.vofo[ci(G, i, nj)], Uc.vofS[ci(G, i, nj)]);
            m1 = vs * r * UDX(i);
            k1 = vs * 0.5 * (UU(i, j) + UU(i, j - 1)) * r * UDX(i);
        }
    } else if (j == 1) {
        if (Uc.aS[gi(G, i, j)] >= EPSI) {
            const double r = rho_mix(G, Uc.vofo[ci(G, i, j)], Uc.vofS[ci(G, i, j)]);
            m1 = vs * r * UDX(i);
            k1 = vs * 0.5 * (UU(i, j) + UU(i, j - 1)) * r * UDX(i);
        }
    } else {
        const double Sx = Vc.SxE[gi(G, i, j - 1)];
        double eta = fabs(Uc.fcnx[gi(G, i, j - 1)] + PDX(i) / 2.0 - Vc.ccx[ci(G, i, j - 1)]) / Sx;
        if (fabs(eta) >= 1.0) eta = 0.5;
        const double w = (1.0 - eta) * VV(i, j - 1) + eta * VV(i + 1, j - 1);
        double p, m;
        upwind(w, p, m);
        const size_t ca = ci(G, i, j - 1), cb = ci(G, i, j);
        const double fa = Uc.vofo[ca], sa = Uc.vofS[ca], fb = Uc.vofo[cb], sb = Uc.vofS[cb];
        if (Uc.aS[gi(G, i, j)] >= EPSI) {
            const double ra = rho_mix(G, fa, sa);
            const double rb = rho_mix(G, fb, sb);
            m1 = (p * ra + m * rb) * UDX(i);
            k1 = (p * UU(i, j - 1) * ra + m * UU(i, j) * rb) * UDX(i);
        }
    }
    // V component
    vs = 0.5 * (VV(i, j) + VV(i, j - 1));
    double w2 = 0.0;
    m2 = 0.0;
    if (j >= nj) {
        if (Vc.aN[gi(G, i, nj)] >= EPSI) w2 = mix_rho(G, G.pvof[ci(G, i, nj)], G.pvofS[ci(G, i, nj)]);
        m2 = vs * w2 * VDX(i);
    } else if (j == 1) {
        if (Vc.aS[gi(G, i, j)] >= EPSI) w2 = mix_rho(G, G.pvof[ci(G, i, j)], G.pvofS[ci(G, i, j)]);
        m2 = vs * w2 * VDX(i);
    } else {
        // operands of every branch below, loaded before the first test
        const size_t c = ci(G, i, j);
        const double as = Vc.aS[gi(G, i, j)], eta = Vc.EtaN[gi(G, i, j - 1)], al = Vc.AlN[gi(G, i, j - 1)];
        const double pv = G.pvof[c], pvs = G.pvofS[c];
        if (as >= EPSI) {
            double w = (1.0 - eta) * VV(i, j - 1) + eta * VV(i, j);
            w = ((w - vb) * al + vb);
            if (pv > 1.0 - EPSI || pv < EPSI) {
                w2 = rho_mix(G, pv, pvs);
            } else if (pvs < EPSI) {
                const double volf =
                    pfrac(G.pnx[c], G.pny[c], G.pphi[c] + w * dt / 2.0 * G.pny[c], PDX(i), fabs(w * dt));
                w2 = (volf * G.rl / G.rr + (1.0 - volf - 0.0) * G.rg / G.rr);
            } else {
                double vols, volf;
                cut_cell_fractions(G.pnxS[c], G.pnyS[c], G.pphiS[c] + w * dt / 2.0 * G.pnyS[c], G.pnx[c], G.pny[c],
                                   G.pphi[c] + w * dt / 2.0 * G.pny[c], PDX(i), fabs(w * dt), vols, volf);
                w2 = (volf * G.rl / G.rr + (1.0 - volf - vols) * G.rg / G.rr);
            }
            m2 = w * w2 * VDX(i);
        }
    }
    // density-weighted V momentum flux
    if (j >= nj || j == 1) {
        const double w = 0.5 * (VV(i, j) + VV(i, j - 1));
        k2 = sq(w) * w2 * VDX(i);
    } else {
        const double eta = Vc.EtaN[gi(G, i, j - 1)];
        double w = (1.0 - eta) * VV(i, j - 1) + eta * VV(i, j);
        w = ((w - vb) * Vc.AlN[gi(G, i, j - 1)] + vb);
        double p, m;
        upwind(w, p, m);
        k2 = (p * VV(i, j - 1) + m * VV(i, j)) * w2 * VDX(i);
    }
}

// ---- viscous fluxes of one face (DiffusiveFlux :2031-2125) -------------------

__device__ void vflux_ew(const PredG& G, int i, int j, bool newvof, double& d1, double& d2, double& e1, double& e2,
                         double& x2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const double* uv = newvof ? Uc.vofn : Uc.vofo;
    const double* vv = newvof ? Vc.vofn : Vc.vofo;
    const double tol = 1.0e-24, Rey = G.Re;
    const int ni = G.ni;
    e1 = e2 = x2 = 0.0;
    if (i == 1) {
        d1 = edge_mix(G, uv[ci(G, 1, j)], Uc.vofS[ci(G, 1, j)]) * Uc.aW[gi(G, 1, j)] * UDY(j) / PDX(1) / Rey;
        d2 = edge_mix(G, vv[ci(G, 1, j)], Vc.vofS[ci(G, 1, j)]) * Vc.aW[gi(G, 1, j)] * VDY(j) / PDX(1) / Rey;
        return;
    }
    if (i == ni + 1) {
        d1 = edge_mix(G, uv[ci(G, ni, j)], Uc.vofS[ci(G, ni, j)]) / Rey * Uc.aE[gi(G, ni, j)] * UDY(j) / PDX(ni);
        d2 = edge_mix(G, vv[ci(G, ni, j)], Vc.vofS[ci(G, ni, j)]) / Rey * Vc.aE[gi(G, ni, j)] * VDY(j) / PDX(ni);
        return;
    }
    double Sx = Uc.SxE[gi(G, i - 1, j)];
    double Sy = Uc.ccy[ci(G, i, j)] - Uc.ccy[ci(G, i - 1, j)];
    double vf = 0.5 * (uv[ci(G, i, j)] / (1.0 - Uc.vofS[ci(G, i, j)] + tol) +
                       uv[ci(G, i - 1, j)] / (1.0 - Uc.vofS[ci(G, i - 1, j)] + tol));
    const double dal1 = Uc.DAlE[gi(G, i - 1, j)];
    d1 = nu_face(G, vf) / Rey * Uc.aW[gi(G, i, j)] * UDY(j) / Sx;
    if (fabs(Sy) >= 1.0e-3 * UDY(j)) e1 = nu_face(G, vf) / Rey * Uc.aW[gi(G, i, j)] * dal1 * UDY(j);
    Sx = Vc.SxE[gi(G, i - 1, j)];
    Sy = Vc.ccy[ci(G, i, j)] - Vc.ccy[ci(G, i - 1, j)];
    vf = 0.5 * (vv[ci(G, i, j)] / (1.0 - Vc.vofS[ci(G, i, j)] + tol) +
                vv[ci(G, i - 1, j)] / (1.0 - Vc.vofS[ci(G, i - 1, j)] + tol));
    const double dal2 = Vc.DAlE[gi(G, i - 1, j)];
    d2 = nu_face(G, vf) / Rey * Vc.aW[gi(G, i, j)] * VDY(j) / Sx;
    if (fabs(Sy) >= 1.0e-3 * VDY(j)) {
        e2 = nu_face(G, vf) / Rey * Vc.aW[gi(G, i, j)] * dal2 * VDY(j);
        x2 = G.vb * e2;
    }
}

__device__ void vflux_ns(const PredG& G, int i, int j, bool newvof, double& d1, double& d2, double& e1, double& e2,
                         double& x2) {
    const StagV& Uc = G.U;
    const StagV& Vc = G.V;
    const double* uv = newvof ? Uc.vofn : Uc.vofo;
    const double* vv = newvof ? Vc.vofn : Vc.vofo;
    const double tol = 1.0e-24, Rey = G.Re;
    const int nj = G.nj;
    e1 = e2 = x2 = 0.0;
    if (j == 1) {
        d1 = edge_mix(G, uv[ci(G, i, 1)], Uc.vofS[ci(G, i, 1)]) / Rey * Uc.aS[gi(G, i, 1)] * UDX(i) / PDY(1);
        d2 = edge_mix(G, vv[ci(G, i, 1)], Vc.vofS[ci(G, i, 1)]) / Rey * Vc.aS[gi(G, i, 1)] * VDX(i) / PDY(1);
        return;
    }
    if (j == nj + 1) {
        d1 = edge_mix(G, uv[ci(G, i, nj)], Uc.vofS[ci(G, i, nj)]) / Rey * Uc.aN[gi(G, i, nj)] * UDX(i) / PDY(nj);
        d2 = edge_mix(G, vv[ci(G, i, nj)], Vc.vofS[ci(G, i, nj)]) / Rey * Vc.aN[gi(G, i, nj)] * VDX(i) / PDY(nj);
        return;
    }
    double Sx = Uc.ccx[ci(G, i, j)] - Uc.ccx[ci(G, i, j - 1)];
    double Sy = Uc.SyN[gi(G, i, j - 1)];
    double vf = 0.5 * (uv[ci(G, i, j - 1)] / (1.0 - Uc.vofS[ci(G, i, j - 1)] + tol) +
                       uv[ci(G, i, j)] / (1.0 - Uc.vofS[ci(G, i, j)] + tol));
    const double dal1 = Uc.DAlN[gi(G, i, j - 1)];
    d1 = nu_face(G, vf) / Rey * Uc.aS[gi(G, i, j)] * UDX(i) / Sy;
    if (fabs(Sx) >= 1.0e-3 * UDX(i)) e1 = nu_face(G, vf) / Rey * Uc.aS[gi(G, i, j)] * dal1 * UDX(i);
    Sx = Vc.ccx[ci(G, i, j)] - Vc.ccx[ci(G, i, j - 1)];
    Sy = Vc.SyN[gi(G, i, j - 1)];
    vf = 0.5 * (vv[ci(G, i, j - 1)] / (1.0 - Vc.vofS[ci(G, i, j - 1)] + tol) +
                vv[ci(G, i, j)] / (1.0 - Vc.vofS[ci(G, i, j)] + tol));
    const double dal2 = Vc.DAlN[gi(G, i, j - 1)];
    d2 = nu_face(G, vf) / Rey * Vc.aS[gi(G, i, j)] * VDX(i) / Sy;
    if (fabs(Sx) >= 1.0e-3 * VDX(i)) {
        e2 = nu_face(G, vf) / Rey * Vc.aS[gi(G, i, j)] * dal2 * VDX(i);
        x2 = G.vb * e2;
    }
}

// ---- kernels ----------------------------------------------------------------------
//
// The face fluxes are not stored in memory: each of the three cell kernels
// computes the faces its tile of FTW x FTH U/V cells reads into LDS, then
// runs its cell update out of LDS.  A tile reads the west and south faces of
// its cells -- face point (i,j) of cell (i,j), one per thread -- and the east
// faces of its last column and north faces of its last row, a halo of
// FTW + FTH faces computed by the first two wavefronts in a second round.

constexpr int FTW = 64, FTH = 8, FNT = FTW * FTH;

// N face quantities of a tile: e[k][lj][li] is EW face (i0+li+1, j0+lj+1),
// li = 0..FTW; n[k][lj][li] is NS face (i0+li+1, j0+lj+1), lj = 0..FTH
template <int N>
struct TilePlanes {
    double e[N][FTH][FTW + 1];
    double n[N][FTH + 1][FTW];
};

// calls ew(i, j, lj, li) for every EW face of the tile that exists
// (j <= nj, i <= ni+1) and ns(i, j, lj, li) for every NS face (i <= ni,
// j <= nj+1); each is called from one place, so its code is emitted once.
// Ends with a barrier.
template <class EW, class NS>
__device__ __forceinline__ void tile_faces(const PredG& G, int i0, int j0, EW&& ew, NS&& ns) {
    const int t = threadIdx.x + threadIdx.y * FTW;
#pragma unroll 1
    for (int r = 0; r < 2; ++r) {
        int li, lj;
        bool e, n;
        if (r == 0) {
            li = threadIdx.x;
            lj = threadIdx.y;
            e = n = true;
        } else if (t < FTW) {
            li = t;  // north faces of the last row
            lj = FTH;
            e = false;
            n = true;
        } else if (t < FTW + FTH) {
            li = FTW;  // east faces of the last column
            lj = t - FTW;
            e = true;
            n = false;
        } else {
            break;
        }
        const int i = i0 + li + 1, j = j0 + lj + 1;
        if (e && j <= G.nj && i <= G.ni + 1) ew(i, j, lj, li);
        if (n && i <= G.ni && j <= G.nj + 1) ns(i, j, lj, li);
    }
    __syncthreads();
}

// the four faces of cell (i0+li+1, j0+lj+1) in plane k
struct Quad {
    double e, w, n, s;
};
template <int N>
__device__ __forceinline__ Quad quad(const TilePlanes<N>& P, int k, int li, int lj) {
    return Quad{P.e[k][lj][li + 1], P.e[k][lj][li], P.n[k][lj + 1][li], P.n[k][lj][li]};
}

struct Comp {  // one momentum component
    const StagV* T;
    const double* w;  // TVar%u or TVar%v
    int imax, jmax;
};

__device__ __forceinline__ Comp comp_of(const PredG& G, int c) {
    return c == 0 ? Comp{&G.U, G.u, G.ni - 1, G.nj} : Comp{&G.V, G.v, G.ni, G.nj - 1};
}

// convective part of the assembly (:175-229 / :320-385): FluxDiv before
// mixing into fd[c]; Flux_n1 updated on active cells.  Faces: CFlux,
// MaFlux and ConFlux (planes 0-1, 2-3, 4-5; U then V) at time n.
__global__ void __launch_bounds__(FNT) k_conv(PredG G, double* __restrict__ fn1u, double* __restrict__ fn1v,
                                              double* __restrict__ fdu, double* __restrict__ fdv) {
    __shared__ TilePlanes<6> P;
    int bx, by;  // the workgroup's tile, XCD-aware
    xcd_tile(bx, by);
    const int i0 = bx * FTW, j0 = by * FTH;
    tile_faces(
        G, i0, j0,
        [&](int i, int j, int lj, int li) {
            double f1, f2, m1, m2, k1, k2;
            cflux_ew(G, i, j, f1, f2);
            dflux_ew(G, i, j, m1, m2, k1, k2);
            P.e[0][lj][li] = f1;
            P.e[1][lj][li] = f2;
            P.e[2][lj][li] = m1;
            P.e[3][lj][li] = m2;
            P.e[4][lj][li] = k1;
            P.e[5][lj][li] = k2;
        },
        [&](int i, int j, int lj, int li) {
            double f1, f2, m1, m2, k1, k2;
            cflux_ns(G, i, j, f1, f2);
            dflux_ns(G, i, j, m1, m2, k1, k2);
            P.n[0][lj][li] = f1;
            P.n[1][lj][li] = f2;
            P.n[2][lj][li] = m1;
            P.n[3][lj][li] = m2;
            P.n[4][lj][li] = k1;
            P.n[5][lj][li] = k2;
        });
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = i0 + li + 1, j = j0 + lj + 1;
    for (int cp = 0; cp < 2; ++cp) {
        const Comp K = comp_of(G, cp);
        if (i > K.imax || j > K.jmax) continue;
        const StagV& T = *K.T;
        const size_t c = ci(G, i, j);
        double* fd = cp == 0 ? fdu : fdv;
        double* fn1 = cp == 0 ? fn1u : fn1v;
        // every operand of the cell, loaded before the activity test
        const size_t g = gi(G, i, j);
        const bool act = T.act[c] != 0;
        const double vf = T.vofo[c], vs = T.vofS[c], wc = K.w[g], f1o = fn1[g];
        const double aw = T.aW[g], ae = T.aE[g], as = T.aS[g], an = T.aN[g];
        if (!act) {
            fd[c] = 0.0;
            continue;
        }
        const double hx = cp == 0 ? UDX(i) : VDX(i), hy = cp == 0 ? UDY(j) : VDY(j);
        const Quad cf = quad(P, cp, li, lj), ma = quad(P, 2 + cp, li, lj), co = quad(P, 4 + cp, li, lj);
        const double Fe = co.e, Fw = co.w, Fn = co.n, Fs = co.s;
        const double RoCeO = (1.0 - vf - vs) * G.rg / G.rr + vf * G.rl / G.rr;
        const double RoCeN = RoCeO - G.dt * (ma.e - ma.w + ma.n - ma.s) / (hx * hy);
        const double lo = (1.0 + EPSI - vs) * G.rg / G.rr, hi = (1.0 - EPSI - vs) * G.rl / G.rr;
        double fluxn0;
        if (RoCeN > lo && RoCeN < hi && RoCeO > lo && RoCeO < hi && vs < EPSI)
            fluxn0 = (Fe - Fw + Fn - Fs) / RoCeN + wc * (1.0 - RoCeO / RoCeN) / G.dt * hx * hy;
        else
            fluxn0 = (cf.e - cf.w + cf.n - cf.s) / (1.0 - vs);
        double v0 = G.first ? fluxn0 : 1.5 * fluxn0 - 0.5 * f1o;
        int closed = 0;
        if (aw < EPSI) closed++;
        if (ae < EPSI) closed++;
        if (as < EPSI) closed++;
        if (an < EPSI) closed++;
        if (closed >= 3) v0 = 0.0;
        fd[c] = v0 - wc * (1.0 - vs) / (1.0 - vs) * hx * hy / G.dt;
        fn1[g] = fluxn0;
    }
}

// whether cell (si,sj) of component K is a mixing source (inside the grid,
// vofS >= 0.5 + EPSI and active; Mixing_Procedure :2139-2141); vs, act are
// its vofS and activity
__device__ __forceinline__ bool is_mix_source(const Comp& K, int si, int sj, double vs, int act) {
    return si >= 1 && si <= K.imax && sj >= 1 && sj <= K.jmax && vs >= 0.5 + EPSI && act;
}

// mixing of source cell (si,sj) of component K (Mixing_Procedure :2139-2171),
// a mixing source: Fm[k], 1/Vt targets
__device__ void mix_source(const PredG& G, const Comp& K, const double* f, int si, int sj, double Fm[3], double Vt[3],
                           int tx[3], int ty[3], bool ok[3]) {
    const StagV& T = *K.T;
    const size_t c = ci(G, si, sj);
    const int sx = (int)fsign(1.0, T.nxS[c]), sy = (int)fsign(1.0, T.nyS[c]);
    tx[0] = si + sx; ty[0] = sj;
    tx[1] = si;      ty[1] = sj + sy;
    tx[2] = si + sx; ty[2] = sj + sy;
    double Be[3];
    const double vs = T.vofS[c];
    for (int k = 0; k < 3; ++k) {
        ok[k] = tx[k] >= 1 && tx[k] <= K.imax && ty[k] >= 1 && ty[k] <= K.jmax;
        if (ok[k]) {
            Vt[k] = 1.0 - T.vofS[ci(G, tx[k], ty[k])];
            Fm[k] = (1.0 - vs) * Vt[k] * (f[ci(G, tx[k], ty[k])] - f[c]);
            Be[k] = k == 0 ? sq(T.nxS[c]) * pow5(Vt[k]) : (k == 1 ? sq(T.nyS[c]) * pow5(Vt[k])
                                                                  : fabs(T.nxS[c] * T.nyS[c]) * pow5(Vt[k]));
        } else {
            Fm[k] = 0.0;
            Be[k] = 0.0;
            Vt[k] = 1.0;
        }
    }
    // a fully solid target (Vt = 0) or a source with no fluid target (all
    // Be = 0) is 0/0 in the reference's expression: its share is 0 here and
    // the target is skipped; other cases are the reference's expression
    const double temp = (Be[0] + Be[1]) + Be[2];
    for (int k = 0; k < 3; ++k) {
        if (temp > 0.0 && Vt[k] > 0.0) {
            Be[k] = Be[k] / temp;
            Fm[k] = Fm[k] * Be[k] / (Be[k] * (1.0 - vs) + Vt[k]);
        } else {
            Fm[k] = 0.0;
            ok[k] = false;
        }
    }
}

// mixing (gathered in the reference's scatter order) + explicit viscous
// half + moving-wall friction (:231-249, :387-420).  fpre: FluxDiv before
// mixing; out: final FluxDiv.  Faces: DifFlux (planes 0-1, U then V) and
// the V component's ExEDFlux (plane 2) with the time-n volume fractions.
__global__ void __launch_bounds__(FNT) k_mix_visc(PredG G, const double* __restrict__ fpre_u,
                                                  const double* __restrict__ fpre_v, double* __restrict__ out_u,
                                                  double* __restrict__ out_v) {
    __shared__ TilePlanes<3> P;
    int bx, by;  // the workgroup's tile, XCD-aware
    xcd_tile(bx, by);
    const int i0 = bx * FTW, j0 = by * FTH;
    tile_faces(
        G, i0, j0,
        [&](int i, int j, int lj, int li) {
            double d1, d2, e1, e2, x2;
            vflux_ew(G, i, j, false, d1, d2, e1, e2, x2);
            P.e[0][lj][li] = d1;
            P.e[1][lj][li] = d2;
            P.e[2][lj][li] = x2;
        },
        [&](int i, int j, int lj, int li) {
            double d1, d2, e1, e2, x2;
            vflux_ns(G, i, j, false, d1, d2, e1, e2, x2);
            P.n[0][lj][li] = d1;
            P.n[1][lj][li] = d2;
            P.n[2][lj][li] = x2;
        });
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = i0 + li + 1, j = j0 + lj + 1;
    for (int cp = 0; cp < 2; ++cp) {
        const Comp K = comp_of(G, cp);
        if (i > K.imax || j > K.jmax) continue;
        const StagV& T = *K.T;
        const double* f = cp == 0 ? fpre_u : fpre_v;
        const size_t c = ci(G, i, j);
        double val = f[c];
        // the mixing sources of the 3x3 window as bits 3a+b of source
        // (i-1+a, j-1+b), their solid fractions and activity loaded
        // together (indices clamped to the grid; the range is in the test)
        unsigned src = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const int si = i - 1 + a, sj = j - 1 + b;
                const size_t cw = ci(G, min(max(si, 1), K.imax), min(max(sj, 1), K.jmax));
                if (is_mix_source(K, si, sj, T.vofS[cw], T.act[cw])) src |= 1u << (3 * a + b);
            }
        // sources (si,sj) in loop order i outer, j inner: ascending bits
        for (; src; src &= src - 1u) {
            const int bit = __builtin_ctz(src);
            const int si = i - 1 + bit / 3, sj = j - 1 + bit % 3;
            {
                double Fm[3], Vt[3];
                int tx[3], ty[3];
                bool ok[3];
                mix_source(G, K, f, si, sj, Fm, Vt, tx, ty, ok);
                if (si == i && sj == j) {
                    val = f[c] + 1.0 / (1.0 - T.vofS[c]) * ((Fm[0] + Fm[1]) + Fm[2]);
                    continue;
                }
                for (int k = 0; k < 3; ++k)
                    if (ok[k] && tx[k] == i && ty[k] == j) val = val - 1.0 / Vt[k] * Fm[k];
            }
        }
        if (T.act[c]) {
            const Quad d = quad(P, cp, li, lj);
            const double* w = K.w;
            double ExDe = 0.0, ExDw = 0.0, ExDn = 0.0, ExDs = 0.0;
            if (cp == 1) {
                const Quad x = quad(P, 2, li, lj);
                ExDe = x.e;
                ExDw = x.w;
                ExDn = x.n;
                ExDs = x.s;
            }
            const double vs = T.vofS[c];
#define W(a, b) w[gi(G, a, b)]
            val = val -
                  BETAVIS * ((d.e * (W(i + 1, j) - W(i, j)) - d.w * (W(i, j) - W(i - 1, j))) -
                             (d.n * (W(i, j + 1) - W(i, j)) - d.s * (W(i, j) - W(i, j - 1)))) /
                      (1.0 - vs) +
                  (-ExDe + ExDw - ExDn + ExDs) / (1.0 - vs);
#undef W
            if (cp == 1 && vs > EPSI && vs < 1.0 - EPSI) {
                const double vf = T.vofn[c];
                val = val - ((1.0 - vf / (1.0 - vs)) * G.nug / G.nur + vf / (1.0 - vs) * G.nul / G.nur) / G.Re *
                                T.WlLh[c] / T.delh[c] * G.vb;
            }
        }
        (cp == 0 ? out_u : out_v)[c] = val;
    }
}

// momentum rows, right-hand sides and Dp (SetMatrix :639-710, SetVectors
// :797-822) for both components; rows in the matr layout (5 planes).
// Faces: DifFlux (planes 0-1) and EDFlux (planes 2-3), U then V, with the
// volume fractions after the transport.
__global__ void __launch_bounds__(FNT) k_rows(PredG G, const double* __restrict__ fdu,
                                              const double* __restrict__ fdv, double* __restrict__ mu,
                                              double* __restrict__ ru, double* __restrict__ mv,
                                              double* __restrict__ rv, double* __restrict__ DpU,
                                              double* __restrict__ DpV) {
    __shared__ TilePlanes<4> P;
    int bx, by;  // the workgroup's tile, XCD-aware
    xcd_tile(bx, by);
    const int i0 = bx * FTW, j0 = by * FTH;
    tile_faces(
        G, i0, j0,
        [&](int i, int j, int lj, int li) {
            double d1, d2, e1, e2, x2;
            vflux_ew(G, i, j, true, d1, d2, e1, e2, x2);
            P.e[0][lj][li] = d1;
            P.e[1][lj][li] = d2;
            P.e[2][lj][li] = e1;
            P.e[3][lj][li] = e2;
        },
        [&](int i, int j, int lj, int li) {
            double d1, d2, e1, e2, x2;
            vflux_ns(G, i, j, true, d1, d2, e1, e2, x2);
            P.n[0][lj][li] = d1;
            P.n[1][lj][li] = d2;
            P.n[2][lj][li] = e1;
            P.n[3][lj][li] = e2;
        });
    const int li = threadIdx.x, lj = threadIdx.y;
    const int i = i0 + li + 1, j = j0 + lj + 1;
    const size_t n = (size_t)G.ni * G.nj;
    for (int cp = 0; cp < 2; ++cp) {
        const Comp K = comp_of(G, cp);
        if (i > G.ni || j > G.nj) continue;
        const StagV& T = *K.T;
        const size_t c = ci(G, i, j);
        double* m = cp == 0 ? mu : mv;
        double* r = cp == 0 ? ru : rv;
        bool act = false;
        double vs = 0.0, etaE = 0.0, etaW = 0.0, etaN = 0.0, etaS = 0.0, vf = 0.0, wl = 0.0, dh = 0.0, fdc = 0.0;
        bool actS = false, actW = false, actE = false, actN = false;
        if (i <= K.imax && j <= K.jmax) {
            // every operand of the row, loaded before the activity test
            act = T.act[c] != 0;
            vs = T.vofS[c];
            etaE = T.EtaE[gi(G, i, j)];
            etaW = T.EtaE[gi(G, i - 1, j)];
            etaN = T.EtaN[gi(G, i, j)];
            etaS = T.EtaN[gi(G, i, j - 1)];
            vf = T.vofn[c];
            wl = T.WlLh[c];
            dh = T.delh[c];
            fdc = (cp == 0 ? fdu : fdv)[c];
            actS = T.act[ci(G, i, max(j - 1, 1))] != 0;
            actW = T.act[ci(G, max(i - 1, 1), j)] != 0;
            actE = T.act[ci(G, min(i + 1, G.ni), j)] != 0;
            actN = T.act[ci(G, i, min(j + 1, G.nj))] != 0;
        }
        if (!act) {
            for (int k = 0; k < 5; ++k) m[k * n + c] = 0.0;
            r[c] = 0.0;
            continue;
        }
        const int iu = cp == 0, iv = cp == 1;
        const double hx = iu ? UDX(i) : VDX(i), hy = iu ? UDY(j) : VDY(j);
        const Quad d = quad(P, cp, li, lj), e = quad(P, 2 + cp, li, lj);
        const double VofAF = 1.0 - vs;
        double De = (1.0 - BETAVIS) * d.e / VofAF;
        double Dw = (1.0 - BETAVIS) * d.w / VofAF;
        double Dn = (1.0 - BETAVIS) * d.n / VofAF;
        const double Ds = (1.0 - BETAVIS) * d.s / VofAF;
        const double Fep = e.e * (1.0 - etaE) / VofAF;
        const double Fem = e.e * etaE / VofAF;
        const double Fwp = e.w * etaW / VofAF;
        const double Fwm = e.w * (1.0 - etaW) / VofAF;
        const double Fnp = e.n * (1.0 - etaN) / VofAF;
        const double Fnm = e.n * etaN / VofAF;
        const double Fsp = e.s * etaS / VofAF;
        const double Fsm = e.s * (1.0 - etaS) / VofAF;
        if (iu) {
            if (j == G.nj) Dn = 0.0;
            if (i == G.ni - iu) De = 0.0;
            if (i == 1) Dw = 0.0;
        } else {
            if (j == G.nj - iv) Dn = 0.0;
        }
        double aE = De, aW = Dw, aN = Dn, aS = Ds;
        const double cwe1 = aW, cwe2 = aE, csn1 = aS, csn2 = aN;
        double aP = hx * hy / G.dt + aE + aW + aN + aS +
                    ((1.0 - vf / (1.0 - vs)) * G.nug / G.nur + vf / (1.0 - vs) * G.nul / G.nur) / G.Re * wl / dh /
                        VofAF;
        aP = aP + Fep - Fwp + Fnp - Fsp;
        aE = aE - Fem;
        aW = aW + Fwm;
        aN = aN - Fnm;
        aS = aS + Fsm;
        (cp == 0 ? DpU : DpV)[gi(G, i, j)] = hx * hy / (aP - aE - aW - aN - aS);
        m[0 * n + c] = (j > 1 && actS) ? -aS : 0.0;
        m[1 * n + c] = (i > 1 && actW) ? -aW : 0.0;
        m[2 * n + c] = aP;
        m[3 * n + c] = (i < G.ni && actE) ? -aE : 0.0;
        m[4 * n + c] = (j < G.nj && actN) ? -aN : 0.0;
        double b = 0.0;
        b = b - fdc;
        if (iv) b = b - 1.0 / (G.Fr * G.Fr) * hx * hy;
        const double* uu = G.u;
        const double* vv = G.v;
        if (i == 1) b = b + cwe1 * ((double)iu * uu[gi(G, i - 1, j)] + (double)iv * vv[gi(G, i - 1, j)]);
        if (i == G.ni - iu) b = b + cwe2 * ((double)iu * uu[gi(G, i + 1, j)] + (double)iv * vv[gi(G, i + 1, j)]);
        if (j == 1) b = b + csn1 * ((double)iu * uu[gi(G, i, j - 1)] + (double)iv * vv[gi(G, i, j - 1)]);
        // SetVectors reads CSN(j,2) (:822); every CSN(.,2) is the zeroed top-row
        // Dn, so the term adds 0 times the ghost velocity
        if (j == G.nj - iv) b = b + 0.0 * csn2 * ((double)iu * uu[gi(G, i, j + 1)] + (double)iv * vv[gi(G, i, j + 1)]);
        r[c] = b;
    }
}

// Pred%u / Pred%v = solution on the rows' cells, 0 on inactive ones
// (DeltaGetValues :847-885, then + dt*GradP with GradP = 0)
__global__ void k_scatter(int ni, int nj, const int* __restrict__ actu, const int* __restrict__ actv,
                          const double* __restrict__ xu, const double* __restrict__ xv, double* __restrict__ up,
                          double* __restrict__ vp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > ni || j > nj) return;
    const size_t c = (size_t)(i - 1) + (size_t)(j - 1) * ni, g = (size_t)i + (size_t)j * (ni + 2);
    if (i <= ni - 1) up[g] = (actu[c] ? xu[c] : 0.0) + 0.0;
    if (j <= nj - 1) vp[g] = (actv[c] ? xv[c] : 0.0) + 0.0;
}

// ghost copies of PU%Dp / PV%Dp (:530-540; every ring value comes from the
// interior) and the open top Pred%v(:,nj) = Pred%v(:,nj-1) (:2209-2212)
__global__ void k_finish(int ni, int nj, double* __restrict__ DU, double* __restrict__ DV, double* __restrict__ vp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i > ni + 1 || j > nj + 1) return;
    const size_t g = (size_t)i + (size_t)j * (ni + 2);
    {
        const bool ring = i == 0 || i >= ni || j == 0 || j == nj + 1;
        if (ring) {
            const int si = i == 0 ? 1 : (i >= ni ? ni - 1 : i);
            const int sj = j == 0 ? 1 : (j == nj + 1 ? nj : j);
            DU[g] = DU[(size_t)si + (size_t)sj * (ni + 2)];
        }
    }
    {
        const bool ring = i == 0 || i == ni + 1 || j == 0 || j >= nj;
        if (ring) {
            const int si = i == 0 ? 1 : (i == ni + 1 ? ni : i);
            const int sj = j == 0 ? 1 : (j >= nj ? nj - 1 : j);
            DV[g] = DV[(size_t)si + (size_t)sj * (ni + 2)];
        }
    }
    if (j == nj && i >= 1 && i <= ni) vp[g] = vp[(size_t)i + (size_t)(nj - 1) * (ni + 2)];
}

#undef UU
#undef VV
#undef PDX
#undef PDY
#undef UDX
#undef UDY
#undef VDX
#undef VDY

StagV stag_view(const StagDev& S, const double* vofo, const double* vofn) {
    StagV v;
    v.vofS = S.a[STG_VOFS]; v.nxS = S.a[STG_NXS]; v.nyS = S.a[STG_NYS]; v.phiS = S.a[STG_PHIS];
    v.WlLh = S.a[STG_WLLH]; v.delh = S.a[STG_DELH]; v.ccx = S.a[STG_CCX]; v.ccy = S.a[STG_CCY];
    v.moex = S.moex; v.act = S.act;
    v.aE = S.a[STG_AE]; v.aW = S.a[STG_AW]; v.aN = S.a[STG_AN]; v.aS = S.a[STG_AS];
    v.EtaE = S.a[STG_ETAE]; v.EtaN = S.a[STG_ETAN]; v.AlE = S.a[STG_ALE]; v.AlN = S.a[STG_ALN];
    v.DAlE = S.a[STG_DALE]; v.DAlN = S.a[STG_DALN]; v.SxE = S.a[STG_SXE]; v.SyN = S.a[STG_SYN];
    v.fcex = S.a[STG_FCEX]; v.fcey = S.a[STG_FCEY]; v.fcnx = S.a[STG_FCNX]; v.fcny = S.a[STG_FCNY];
    v.vofo = vofo;
    v.vofn = vofn;
    return v;
}

}  // namespace

int predictor_allocate(Ctx* c) {
    const size_t n = (size_t)c->ni * c->nj, ng = (size_t)(c->ni + 2) * (c->nj + 2);
    PredBuffers& B = c->pred;
    for (int s = 0; s < 2; ++s) {
        StagDev& S = c->stag[s];
        for (int k = 0; k < STG_COUNT; ++k) {
            S.a[k] = (double*)dev_alloc(c, (k < STG_NCELL ? n : ng) * sizeof(double));
            if (!S.a[k]) return SG_ERR_ALLOC;
        }
        S.moex = (int*)dev_alloc(c, n * sizeof(int));
        S.act = (int*)dev_alloc(c, n * sizeof(int));
        if (!S.moex || !S.act) return SG_ERR_ALLOC;
    }
    double** cellv[] = {&B.fdu, &B.fdv, &B.gu, &B.gv, &B.ru, &B.rv, &B.xu, &B.xv};
    for (double** p : cellv) {
        *p = (double*)dev_alloc(c, n * sizeof(double));
        if (!*p) return SG_ERR_ALLOC;
    }
    B.mu = (double*)dev_alloc(c, 5 * n * sizeof(double));
    B.mv = (double*)dev_alloc(c, 5 * n * sizeof(double));
    if (!B.mu || !B.mv) return SG_ERR_ALLOC;
    for (double*& v : B.vec) {
        v = (double*)dev_alloc(c, n * sizeof(double));
        if (!v) return SG_ERR_ALLOC;
    }
    B.partials = (double*)dev_alloc(c, 4 * PredBuffers::BI_BLOCKS * sizeof(double));
    B.sc = (BiScalars*)dev_alloc(c, sizeof(BiScalars));
    if (!B.partials || !B.sc) return SG_ERR_ALLOC;
    if (hipHostMalloc((void**)&B.sc_host, sizeof(BiScalars), hipHostMallocDefault) != hipSuccess) return SG_ERR_ALLOC;
    return SG_OK;
}

// Predictor_UV up to the linear systems: Flux_n1, Dp on the active cells,
// rows and right-hand sides in c->pred.
int predictor_assemble(Ctx* c, double dt, double vb, int first_step) {
    PredG G;
    G.ni = c->ni;
    G.nj = c->nj;
    G.dx = c->dx;
    G.dy = c->dy;
    G.udx = c->udx;
    G.vdy = c->vdy;
    auto fd = [&](int f) { return (const double*)c->field[f]; };
    G.U = stag_view(c->stag[0], fd(SG_F_UVOFO), fd(SG_F_UVOF));
    G.V = stag_view(c->stag[1], fd(SG_F_VVOFO), fd(SG_F_VVOF));
    G.pvof = fd(SG_F_VOFO);
    G.pphi = fd(SG_F_PHIO);
    G.pnx = fd(SG_F_NXO);
    G.pny = fd(SG_F_NYO);
    G.pvofS = fd(SG_F_VOFS);
    G.pnxS = fd(SG_F_NXS);
    G.pnyS = fd(SG_F_NYS);
    G.pphiS = fd(SG_F_PHIS);
    G.u = fd(SG_F_U);
    G.v = fd(SG_F_V);
    G.rl = c->prm.rho_l;
    G.rg = c->prm.rho_g;
    G.rr = c->prm.rho_ref;
    G.nul = c->prm.nu_l;
    G.nug = c->prm.nu_g;
    G.nur = c->prm.nu_ref;
    G.Re = c->prm.Re;
    G.Fr = c->prm.Fr;
    G.dt = dt;
    G.vb = vb;
    G.first = first_step ? 1 : 0;
    PredBuffers& B = c->pred;
    const dim3 blk(FTW, FTH), grd((c->ni + FTW - 1) / FTW, (c->nj + FTH - 1) / FTH);
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_conv, grd, blk, 0, st, G, (double*)c->field[SG_F_FLUXN1U], (double*)c->field[SG_F_FLUXN1V],
                       B.gu, B.gv);
    hipLaunchKernelGGL(k_mix_visc, grd, blk, 0, st, G, B.gu, B.gv, B.fdu, B.fdv);
    hipLaunchKernelGGL(k_rows, grd, blk, 0, st, G, B.fdu, B.fdv, B.mu, B.ru, B.mv, B.rv,
                       (double*)c->field[SG_F_DPU], (double*)c->field[SG_F_DPV]);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int predictor_finish(Ctx* c) {
    const dim3 blk(64, 4);
    const dim3 gc((c->ni + 63) / 64, (c->nj + 3) / 4);
    const dim3 gr((c->ni + 2 + 63) / 64, (c->nj + 2 + 3) / 4);
    hipLaunchKernelGGL(k_scatter, gc, blk, 0, c->stream, c->ni, c->nj, (const int*)c->stag[0].act,
                       (const int*)c->stag[1].act, c->pred.xu, c->pred.xv, (double*)c->field[SG_F_UPRED],
                       (double*)c->field[SG_F_VPRED]);
    hipLaunchKernelGGL(k_finish, gr, blk, 0, c->stream, c->ni, c->nj, (double*)c->field[SG_F_DPU],
                       (double*)c->field[SG_F_DPV], (double*)c->field[SG_F_VPRED]);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace sg
