/*
 * CPU FP64 oracle of the momentum predictor (PredictorUV_mod.f90).
 * TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU-baseline leg
 * of bench.py call it as the checker; the product path never links it.
 *
 * Each routine restates the reference subroutine it names, in its loop
 * order and with its operand order (compiled with -ffp-contract=off).
 * x**2.d0 is x*x; x**5.d0 is ((x*x)*(x*x))*x (the reference calls pow).
 * MIN/MAX follow gfortran's intrinsics (fmin/fmax: a NaN argument is
 * ignored).  Deliberate departures, each at its site:
 *   - FaceDensityFlux's two-phase branch reads `vols` before any
 *     assignment in that branch (:1924,2009); it is taken as 0 (no solid).
 *   - SetVectors indexes CSN(j,2) for CSN(i,2) (:822); every CSN(.,2) entry
 *     is 0 (the top-row Dn is zeroed before CSN is filled), so the term is
 *     0, which is what is computed here.
 *   - the bottom-row U flux `alr*ur*+alr*ul` (:1322) is parsed as gfortran
 *     does, ((alr*ur)*alr)*ul.
 *
 * Layout: cell arrays a[(i-1)+(j-1)*ni]; ghosted arrays a[i+j*(ni+2)].
 * Face arrays EW (i=1..ni+1, j=1..nj) and NS (i=1..ni, j=1..nj+1) are
 * stored in ghosted arrays at (i,j).
 *
 * Pinned by closed forms (tests/test_oracle_predictor.py): steady uniform
 * stream, gravity increment and Dp = dt at rest, steady Couette shear, the
 * Adams-Bashforth history; the limiter against Sweby's superbee table, its
 * TVD region and symmetry; the MUSCL / kinetic face fluxes exact on linear
 * profiles; the mixing of a cell with one target (the volume-weighted mean),
 * its conservation of sum (1 - vofS) flux, and the reference's overwrite of
 * a small cell that an earlier one mixed into; the first and last faces'
 * fluxes (local Lax-Friedrichs on expansions, upwind for the carried
 * component) on stretched grids; a source split equally among three
 * targets; the mass-momentum consistency of the density-based faces (a
 * uniform stream across an inclined interface stays uniform).  Parity
 * unpinned (restated line by line): the swept-region density those faces
 * share (pfrac / CellGeoCal of the swept box), how Beta weighs the
 * targets' fluid volumes (the fifth power), the first NS face's U momentum
 * as gfortran parses it, and the cut-cell interpolation branches
 * (modified_conv and the near-solid branches of high_order_x/y, whose
 * geometry inputs are pinned, not their use).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "sg_oracle.h"

#define EPSI 1.0e-3     /* StateVariables.f90:28 */
#define BETAVIS 0.5     /* StateVariables.f90:28 */
#define TOL14 1.0e-14   /* PredictorUV_mod.f90:16 */

typedef struct {
    const orc_pred_in* in;
    int ni, nj;
} PC;

#define NI (p->ni)
#define NJ (p->nj)
#define CI(i, j) ((size_t)((i)-1) + (size_t)((j)-1) * (size_t)NI)
#define GI(i, j) ((size_t)(i) + (size_t)(j) * (size_t)(NI + 2))
#define U(i, j) (p->in->u[GI(i, j)])
#define V(i, j) (p->in->v[GI(i, j)])
/* grids: P (dx, dy), U (udx, dy), V (dx, vdy) */
#define PDX(i) (p->in->dx[(i)-1])
#define PDY(j) (p->in->dy[(j)-1])
#define UDX(i) (p->in->udx[(i)-1])
#define UDY(j) (p->in->dy[(j)-1])
#define VDX(i) (p->in->dx[(i)-1])
#define VDY(j) (p->in->vdy[(j)-1])

static double sq(double x) { return x * x; }
static double pow5(double x) { return ((x * x) * (x * x)) * x; }
static double fsign(double a, double b) { return copysign(fabs(a), b); }

/* MUSCLLimiter, PredictorUV_mod.f90:2354-2367 (superbee, opt 2) */
static double muscl(double x, double tolim)
{
    return fmax(fmax(0.0, fmin(tolim * x, 1.0)), fmin(x, tolim));
}

/* frac of PredictorUV_mod.f90:2215-2257 (as Clsvof's, vrt = 0 first) */
static double frac(double nx, double ny, double diss, double dx, double dy)
{
    return orc_frac(nx, ny, diss, dx, dy);
}

/* ---- convective fluxes ---------------------------------------------------- */

/* ModifiedConvectiveFlux, :888-997 (first step) */
static void modified_conv(const PC* p, double* F1, double* F2, int idir)
{
    const orc_stag* Uc = &p->in->U;
    const orc_stag* Vc = &p->in->V;
    double vb = p->in->vb;
    for (int i = 1; i <= NI + idir; i++) {
        for (int j = 1; j <= NJ + (1 - idir); j++) {
            if (idir) {
                double uw = 0.5 * (U(i, j) + U(i - 1, j));
                if (i == 1) {
                    F1[GI(i, j)] = sq(uw) * UDY(j) * Uc->aW[GI(i, j)];
                } else if (i >= NI) {
                    F1[GI(i, j)] = sq(uw) * UDY(j) * Uc->aE[GI(i - 1, j)];
                } else {
                    double eta = Uc->EtaE[GI(i - 1, j)];
                    uw = (1.0 - eta) * U(i - 1, j) + eta * U(i, j);
                    double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                    F1[GI(i, j)] = (uwp * U(i - 1, j) + uwn * U(i, j)) * Uc->AlE[GI(i - 1, j)] * UDY(j) *
                                   Uc->aW[GI(i, j)];
                }
                uw = 0.5 * (U(i - 1, j + 1) + U(i - 1, j));
                if (i > NI) {
                    F2[GI(i, j)] = uw * 0.5 * (V(i - 1, j) + V(i, j)) * VDY(j) * Vc->aE[GI(i - 1, j)];
                } else if (i == 1) {
                    F2[GI(i, j)] = uw * 0.5 * (V(i - 1, j) + V(i, j)) * VDY(j) * Vc->aW[GI(i, j)];
                } else {
                    double Sy = Uc->SyN[GI(i - 1, j)];
                    double eta = fabs(Vc->fcey[GI(i - 1, j)] + PDY(j) / 2.0 - Uc->ccy[CI(i - 1, j)]) / Sy;
                    if (fabs(eta) >= 1.0) eta = 0.5;
                    uw = (1.0 - eta) * U(i - 1, j) + eta * U(i - 1, j + 1);
                    double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                    F2[GI(i, j)] = (uwp * V(i - 1, j) + uwn * V(i, j)) * Vc->aW[GI(i, j)] * VDY(j);
                }
            } else {
                double vs = 0.5 * (V(i, j - 1) + V(i + 1, j - 1));
                if (j > NJ) {
                    F1[GI(i, j)] = vs * 0.5 * (U(i, j) + U(i, j - 1)) * UDX(i) * Uc->aN[GI(i, j - 1)];
                } else if (j == 1) {
                    F1[GI(i, j)] = vs * 0.5 * (U(i, j) + U(i, j - 1)) * UDX(i) * Uc->aS[GI(i, j)];
                } else {
                    double Sx = Vc->SxE[GI(i, j - 1)];
                    double eta = fabs(Uc->fcnx[GI(i, j - 1)] + PDX(i) / 2.0 - Vc->ccx[CI(i, j - 1)]) / Sx;
                    if (fabs(eta) >= 1.0) eta = 0.5;
                    vs = (1.0 - eta) * V(i, j - 1) + eta * V(i + 1, j - 1);
                    double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                    F1[GI(i, j)] = (vsp * U(i, j - 1) + vsn * U(i, j)) * Uc->aS[GI(i, j)] * UDX(i);
                }
                vs = 0.5 * (V(i, j) + V(i, j - 1));
                if (j >= NJ) {
                    F2[GI(i, j)] = sq(vs) * VDX(i) * Vc->aN[GI(i, j - 1)];
                } else if (j == 1) {
                    F2[GI(i, j)] = sq(vs) * VDX(i) * Vc->aS[GI(i, j)];
                } else {
                    double eta = Vc->EtaN[GI(i, j - 1)];
                    vs = (1.0 - eta) * V(i, j - 1) + eta * V(i, j);
                    vs = ((vs - vb) * Vc->AlN[GI(i, j - 1)] + vb);
                    double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                    F2[GI(i, j)] = (vsp * V(i, j - 1) + vsn * V(i, j)) * Vc->aS[GI(i, j)] * VDX(i);
                }
            }
        }
    }
}

/* HighOrderConvectiveFluxForXDir, :1186-1305 */
static void high_order_x(const PC* p, double* F1, double* F2)
{
    const orc_stag* Uc = &p->in->U;
    const orc_stag* Vc = &p->in->V;
    const double tol = 1.0e-24;
    for (int j = 1; j <= NJ; j++) {
        double ul = U(0, j), ur = U(1, j);
        double alr = fmax(fabs(ur), fabs(ul));
        F1[GI(1, j)] = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc->aW[GI(1, j)] * UDY(j);
        double vl = V(0, j), vr = V(1, j);
        alr = 0.5 * (U(0, j) + U(0, j + 1));
        F2[GI(1, j)] = 0.5 * (alr * vr + alr * vl - fabs(alr) * (vr - vl)) * Vc->aW[GI(1, j)] * VDY(j);
        for (int i = 2; i <= NI; i++) {
            int ip = i + 1 < NI ? i + 1 : NI;
            double tolim = fmin(2.0 * UDX(i - 1) / PDX(i), 2.0 * UDX(i) / PDX(i));
            double ri = ((U(i - 1, j) - U(i - 2, j)) / PDX(i - 1)) / ((U(i, j) - U(i - 1, j)) / PDX(i));
            double omei = muscl(ri, tolim) * (U(i, j) - U(i - 1, j)) / PDX(i);
            double ri1 = ((U(i, j) - U(i - 1, j)) / PDX(i)) / ((U(i + 1, j) - U(i, j)) / PDX(ip));
            double omei1 = muscl(ri1, tolim) * (U(i + 1, j) - U(i, j)) / PDX(ip);
            ul = U(i - 1, j) + 0.5 * (PDX(i)) * omei;
            ur = U(i, j) - 0.5 * (PDX(i)) * omei1;
            alr = fmax(fabs(ur), fabs(ul));
            F1[GI(i, j)] = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc->aW[GI(i, j)] * UDY(j);
            if (i > 2 && i < NI) {
                if (Uc->vofS[CI(i - 2, j)] > EPSI || Uc->vofS[CI(i - 1, j)] > EPSI || Uc->vofS[CI(i, j)] > EPSI ||
                    Uc->vofS[CI(i + 1, j)] > EPSI) {
                    double eta = Uc->EtaE[GI(i - 1, j)];
                    double uw = (1 - eta) * U(i - 1, j) + eta * U(i, j);
                    double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                    F1[GI(i, j)] = (uwp * U(i - 1, j) + uwn * U(i, j)) * Uc->AlE[GI(i - 1, j)] * UDY(j) *
                                   Uc->aW[GI(i, j)];
                }
            }
            int im2 = i - 2 > 1 ? i - 2 : 1;
            tolim = fmin(2.0 * VDX(i - 1) / UDX(i - 1), 2.0 * VDX(i) / UDX(i - 1));
            ri = ((V(i - 1, j) - V(i - 2, j)) / UDX(im2)) / ((V(i, j) - V(i - 1, j)) / UDX(i - 1));
            omei = muscl(ri, tolim) * (V(i, j) - V(i - 1, j)) / UDX(i - 1);
            ri1 = ((V(i, j) - V(i - 1, j)) / UDX(i - 1)) / ((V(i + 1, j) - V(i, j)) / UDX(i));
            omei1 = muscl(ri1, tolim) * (V(i + 1, j) - V(i, j)) / UDX(i);
            vl = V(i - 1, j) + 0.5 * (PDX(i - 1)) * omei;
            vr = V(i, j) - 0.50 * (PDX(i)) * omei1;
            alr = 0.5 * (U(i - 1, j) + U(i - 1, j + 1));
            F2[GI(i, j)] = 0.5 * (alr * vr + alr * vl - fabs(alr) * (vr - vl)) * Vc->aW[GI(i, j)] * VDY(j);
            if (i > 2 && i < NI) {
                if (Vc->vofS[CI(i - 2, j)] > EPSI || Vc->vofS[CI(i - 1, j)] > EPSI || Vc->vofS[CI(i, j)] > EPSI ||
                    Vc->vofS[CI(i + 1, j)] > EPSI) {
                    if (Vc->aW[GI(i, j)] < 0.5) {
                        double delhec = fabs(Vc->fcex[GI(i - 1, j)] * Vc->nxS[CI(i - 1, j)] +
                                             Vc->fcey[GI(i - 1, j)] * Vc->nyS[CI(i - 1, j)] + Vc->phiS[CI(i - 1, j)]);
                        double delh, uw;
                        if (Uc->moex[CI(i - 1, j + 1)] != 1 && Uc->vofS[CI(i - 1, j + 1)] < 1.0 - EPSI) {
                            delh = fabs(Uc->ccx[CI(i - 1, j + 1)] * Uc->nxS[CI(i - 1, j + 1)] +
                                        Uc->ccy[CI(i - 1, j + 1)] * Uc->nyS[CI(i - 1, j + 1)] +
                                        Uc->phiS[CI(i - 1, j + 1)]) + tol;
                            uw = U(i - 1, j + 1);
                        } else if (Uc->moex[CI(i - 1, j)] != 1 && Uc->vofS[CI(i - 1, j)] < 1.0 - EPSI) {
                            delh = fabs(Uc->ccx[CI(i - 1, j)] * Uc->nxS[CI(i - 1, j)] +
                                        Uc->ccy[CI(i - 1, j)] * Uc->nyS[CI(i - 1, j)] + Uc->phiS[CI(i - 1, j)]) +
                                   tol;
                            uw = U(i - 1, j);
                        } else {
                            delh = delhec + tol;
                            uw = 0.0;
                        }
                        double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                        F2[GI(i, j)] = (uwp * V(i - 1, j) + uwn * V(i, j)) * delhec / delh * Vc->aW[GI(i, j)] * VDY(j);
                    } else {
                        double Sy = Uc->SyN[GI(i - 1, j)];
                        double eta = fabs(Vc->fcey[GI(i - 1, j)] + VDY(j) / 2.0 - Uc->ccy[CI(i - 1, j)]) / Sy;
                        if (fabs(eta) >= 1.0) eta = 0.5;
                        double uw = (1.0 - eta) * U(i - 1, j) + eta * U(i - 1, j + 1);
                        double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                        F2[GI(i, j)] = (uwp * V(i - 1, j) + uwn * V(i, j)) * Vc->aW[GI(i, j)] * VDY(j);
                    }
                }
            }
        }
        ul = U(NI, j);
        ur = U(NI + 1, j);
        alr = fmax(fabs(ur), fabs(ul));
        F1[GI(NI + 1, j)] = 0.5 * (sq(ur) + sq(ul) - fabs(alr) * (ur - ul)) * Uc->aE[GI(NI, j)] * UDY(j);
        vl = V(NI, j);
        vr = V(NI + 1, j);
        alr = 0.5 * (U(NI, j) + U(NI, j + 1));
        F2[GI(NI + 1, j)] = 0.5 * (alr * vr + alr * vl - fabs(alr) * (vr - vl)) * Vc->aE[GI(NI, j)] * VDY(j);
    }
}

/* HighOrderConvectiveFluxForYDir, :1307-1429 */
static void high_order_y(const PC* p, double* F1, double* F2)
{
    const orc_stag* Uc = &p->in->U;
    const orc_stag* Vc = &p->in->V;
    const double tol = 1.0e-24;
    double vb = p->in->vb;
    for (int i = 1; i <= NI; i++) {
        double ul = U(i, 0), ur = U(i, 1);
        double alr = 0.5 * (V(i, 0) + V(i + 1, 0));
        F1[GI(i, 1)] = 0.5 * (((alr * ur) * alr) * ul - fabs(alr) * (ur - ul)) * Uc->aS[GI(i, 1)] * UDX(i);
        double vl = V(i, 0), vr = V(i, 1);
        alr = fmax(fabs(vl), fabs(vr));
        F2[GI(i, 1)] = 0.5 * (sq(vr) + sq(vl) - fabs(alr) * (vr - vl)) * Vc->aS[GI(i, 1)] * VDX(i);
        for (int j = 2; j <= NJ; j++) {
            int jm2 = j - 2 > 1 ? j - 2 : 1;
            int jp = j + 1 < NJ ? j + 1 : NJ;
            double tolim = fmin(2.0 * UDY(j - 1) / VDY(j - 1), 2.0 * UDY(j) / VDY(j - 1));
            double ri = ((U(i, j - 1) - U(i, j - 2)) / VDY(jm2)) / ((U(i, j) - U(i, j - 1)) / VDY(j - 1));
            double omei = muscl(ri, tolim) * (U(i, j) - U(i, j - 1)) / VDY(j - 1);
            double ri1 = ((U(i, j) - U(i, j - 1)) / VDY(j - 1)) / ((U(i, j + 1) - U(i, j)) / VDY(j));
            double omei1 = muscl(ri1, tolim) * (U(i, j + 1) - U(i, j)) / VDY(j);
            ul = U(i, j - 1) + 0.5 * (PDY(j - 1)) * omei;
            ur = U(i, j) - 0.5 * (PDY(j)) * omei1;
            alr = 0.5 * (V(i, j - 1) + V(i + 1, j - 1));
            F1[GI(i, j)] = 0.5 * (alr * ur + alr * ul - fabs(alr) * (ur - ul)) * Uc->aS[GI(i, j)] * UDX(i);
            if (j > 2 && j < NJ) {
                if (Uc->vofS[CI(i, j - 2)] > EPSI || Uc->vofS[CI(i, j - 1)] > EPSI || Uc->vofS[CI(i, j)] > EPSI ||
                    Uc->vofS[CI(i, j + 1)] > EPSI) {
                    if (Uc->aS[GI(i, j)] < 0.5) {
                        double delhec = fabs(Uc->fcnx[GI(i, j - 1)] * Uc->nxS[CI(i, j - 1)] +
                                             Uc->fcny[GI(i, j - 1)] * Uc->nyS[CI(i, j - 1)] + Uc->phiS[CI(i, j - 1)]);
                        double delh, vs;
                        if (Vc->moex[CI(i, j - 1)] != 1 && Vc->vofS[CI(i, j - 1)] < 1.0 - EPSI) {
                            delh = fabs(Vc->ccx[CI(i, j - 1)] * Vc->nxS[CI(i, j - 1)] +
                                        Vc->ccy[CI(i, j - 1)] * Vc->nyS[CI(i, j - 1)] + Vc->phiS[CI(i, j - 1)]) +
                                   tol;
                            vs = V(i, j - 1);
                        } else if (Vc->moex[CI(i + 1, j - 1)] != 1 && Vc->vofS[CI(i + 1, j - 1)] < 1.0 - EPSI) {
                            delh = fabs(Vc->ccx[CI(i + 1, j - 1)] * Vc->nxS[CI(i + 1, j - 1)] +
                                        Vc->ccy[CI(i + 1, j - 1)] * Vc->nyS[CI(i + 1, j - 1)] +
                                        Vc->phiS[CI(i + 1, j - 1)]) + tol;
                            vs = V(i + 1, j - 1);
                        } else {
                            delh = delhec + tol;
                            vs = 0.0;
                        }
                        vs = vb + (vs - vb) * delhec / delh;
                        double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                        F1[GI(i, j)] = (vsp * U(i, j - 1) + vsn * U(i, j)) * delhec / delh * Uc->aS[GI(i, j)] * UDX(i);
                    } else {
                        double Sx = Vc->SxE[GI(i, j - 1)];
                        double eta = fabs(Uc->fcnx[GI(i, j - 1)] + 0.5 * UDY(j - 1) - Vc->ccx[CI(i, j - 1)]) / Sx;
                        if (fabs(eta) >= 1.0) eta = 0.5;
                        double vs = (1.0 - eta) * V(i, j - 1) + eta * V(i + 1, j - 1);
                        double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                        F1[GI(i, j)] = (vsp * U(i, j - 1) + vsn * U(i, j)) * Uc->aS[GI(i, j)] * UDX(i);
                    }
                }
            }
            tolim = fmin(2.0 * VDY(j - 1) / PDY(j), 2.0 * VDY(j) / PDY(j));
            ri = ((V(i, j - 1) - V(i, j - 2)) / PDY(j - 1)) / ((V(i, j) - V(i, j - 1)) / PDY(j));
            omei = muscl(ri, tolim) * (V(i, j) - V(i, j - 1)) / PDY(j - 1);
            ri1 = ((V(i, j) - V(i, j - 1)) / PDY(j)) / ((V(i, j + 1) - V(i, j)) / PDY(jp));
            omei1 = muscl(ri1, tolim) * (V(i, j + 1) - V(i, j)) / PDY(jp);
            vl = V(i, j - 1) + 0.5 * (PDY(j)) * omei;
            vr = V(i, j) - 0.5 * (PDY(j)) * omei1;
            alr = fmax(fabs(vl), fabs(vr));
            F2[GI(i, j)] = 0.5 * (vr * vr + vl * vl - fabs(alr) * (vr - vl)) * Vc->aS[GI(i, j)] * VDX(i);
            if (j > 2 && j < NJ) {
                if (Vc->vofS[CI(i, j - 2)] > EPSI || Vc->vofS[CI(i, j - 1)] > EPSI || Vc->vofS[CI(i, j)] > EPSI ||
                    Vc->vofS[CI(i, j + 1)] > EPSI) {
                    double eta = Vc->EtaN[GI(i, j - 1)];
                    double vs = (1.0 - eta) * V(i, j - 1) + eta * V(i, j);
                    vs = (vb + (vs - vb) * Vc->AlN[GI(i, j - 1)]);
                    double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                    F2[GI(i, j)] = (vsp * V(i, j - 1) + vsn * V(i, j)) * Vc->aS[GI(i, j)] * VDX(i);
                }
            }
        }
        ul = U(i, NJ);
        ur = U(i, NJ + 1);
        alr = 0.5 * (V(i, NJ) + V(i + 1, NJ));
        F1[GI(i, NJ + 1)] = 0.5 * (alr * ur + alr * ul - fabs(alr) * (ur - ul)) * Uc->aN[GI(i, NJ)] * UDX(i);
        vl = V(i, NJ);
        vr = V(i, NJ + 1);
        alr = fmax(fabs(vl), fabs(vr));
        F2[GI(i, NJ + 1)] = 0.5 * (sq(vr) + sq(vl) - fabs(alr) * (vr - vl)) * Vc->aN[GI(i, NJ)] * VDX(i);
    }
}

/* ---- density-weighted fluxes ------------------------------------------------ */

static double mix_rho(const PC* p, double vf, double vs)
{
    /* (1-vof-vofS)*roa/Roref + vof*row/Roref */
    const orc_pred_in* in = p->in;
    return ((1.0 - vf - vs) * in->rho_g / in->rho_ref + vf * in->rho_l / in->rho_ref);
}
static double rho_mix(const PC* p, double vf, double vs)
{
    /* vof*row/Roref + (1-vof-vofS)*roa/Roref */
    const orc_pred_in* in = p->in;
    return (vf * in->rho_l / in->rho_ref + (1.0 - vf - vs) * in->rho_g / in->rho_ref);
}

/* FaceDensityFlux, :1871-2029 (P cells and U/V vof at time n) */
static void face_density(const PC* p, double* M1, double* M2, double* W1, double* W2, int idir)
{
    const orc_pred_in* in = p->in;
    const orc_stag* Uc = &in->U;
    const orc_stag* Vc = &in->V;
    const double dt = in->dt;
    double vols = 0.0, volf = 0.0;
    for (int i = 1; i <= NI + idir; i++) {
        for (int j = 1; j <= NJ + (1 - idir); j++) {
            if (idir) {
                double uw = 0.5 * (U(i, j) + U(i - 1, j));
                W1[GI(i, j)] = 0.0;
                M1[GI(i, j)] = 0.0;
                if (i == 1) {
                    if (Uc->aW[GI(1, j)] >= EPSI) W1[GI(i, j)] = mix_rho(p, in->pvof[CI(i, j)], in->pvofS[CI(i, j)]);
                    M1[GI(i, j)] = uw * W1[GI(i, j)] * UDY(j);
                } else if (i >= NI) {
                    if (Uc->aE[GI(NI, j)] >= EPSI) W1[GI(i, j)] = mix_rho(p, in->pvof[CI(NI, j)], in->pvofS[CI(NI, j)]);
                    M1[GI(i, j)] = uw * W1[GI(i, j)] * UDY(j);
                } else {
                    if (Uc->aW[GI(i, j)] >= EPSI) {
                        double eta = Uc->EtaE[GI(i - 1, j)];
                        uw = (1.0 - eta) * U(i - 1, j) + eta * U(i, j);
                        size_t c = CI(i, j);
                        if (in->pvof[c] > 1.0 - EPSI || in->pvof[c] < EPSI) {
                            W1[GI(i, j)] = mix_rho(p, in->pvof[c], in->pvofS[c]);
                        } else if (in->pvofS[c] < EPSI) {
                            volf = frac(in->pnx[c], in->pny[c], in->pphi[c] + uw * dt / 2.0 * in->pnx[c],
                                        fabs(uw * dt), PDY(j));
                            W1[GI(i, j)] = (volf * in->rho_l / in->rho_ref + (1.0 - volf - 0.0) * in->rho_g / in->rho_ref);
                        } else {
                            orc_cellgeocal(in->pnxS[c], in->pnyS[c], in->pphiS[c] + uw * dt / 2.0 * in->pnxS[c],
                                           in->pnx[c], in->pny[c], in->pphi[c] + uw * dt / 2.0 * in->pnx[c],
                                           fabs(uw * dt), PDY(j), &vols, &volf);
                            W1[GI(i, j)] = (volf * in->rho_l / in->rho_ref + (1.0 - volf - vols) * in->rho_g / in->rho_ref);
                        }
                    }
                    M1[GI(i, j)] = uw * Uc->AlE[GI(i - 1, j)] * W1[GI(i, j)] * UDY(j);
                }
                uw = 0.5 * (U(i - 1, j + 1) + U(i - 1, j));
                W2[GI(i, j)] = 0.0;
                M2[GI(i, j)] = 0.0;
                if (i > NI) {
                    if (Vc->aE[GI(NI, j)] >= EPSI)
                        M2[GI(i, j)] = uw * rho_mix(p, Vc->vof_old[CI(NI, j)], Vc->vofS[CI(NI, j)]) * VDY(j);
                } else if (i == 1) {
                    if (Vc->aW[GI(1, j)] >= EPSI)
                        M2[GI(i, j)] = uw * rho_mix(p, Vc->vof_old[CI(i, j)], Vc->vofS[CI(i, j)]) * VDY(j);
                } else {
                    if (Vc->aW[GI(i, j)] >= EPSI) {
                        double Sy = Uc->SyN[GI(i - 1, j)];
                        double eta = fabs(Vc->fcey[GI(i - 1, j)] + PDY(j) / 2.0 - Uc->ccy[CI(i - 1, j)]) / Sy;
                        if (fabs(eta) >= 1.0) eta = 0.5;
                        uw = (1.0 - eta) * U(i - 1, j) + eta * U(i - 1, j + 1);
                        double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                        M2[GI(i, j)] = (uwp * rho_mix(p, Vc->vof_old[CI(i - 1, j)], Vc->vofS[CI(i - 1, j)]) +
                                        uwn * rho_mix(p, Vc->vof_old[CI(i, j)], Vc->vofS[CI(i, j)])) *
                                       VDY(j);
                    }
                }
            } else {
                W1[GI(i, j)] = 0.0;
                M1[GI(i, j)] = 0.0;
                double vs = 0.5 * (V(i, j - 1) + V(i + 1, j - 1));
                if (j > NJ) {
                    if (Uc->aN[GI(i, NJ)] >= EPSI)
                        M1[GI(i, j)] = vs * rho_mix(p, Uc->vof_old[CI(i, NJ)], Uc->vofS[CI(i, NJ)]) * UDX(i);
                } else if (j == 1) {
                    if (Uc->aS[GI(i, j)] >= EPSI)
                        M1[GI(i, j)] = vs * rho_mix(p, Uc->vof_old[CI(i, j)], Uc->vofS[CI(i, j)]) * UDX(i);
                } else {
                    if (Uc->aS[GI(i, j)] >= EPSI) {
                        double Sx = Vc->SxE[GI(i, j - 1)];
                        double eta = fabs(Uc->fcnx[GI(i, j - 1)] + PDX(i) / 2.0 - Vc->ccx[CI(i, j - 1)]) / Sx;
                        if (fabs(eta) >= 1.0) eta = 0.5;
                        vs = (1.0 - eta) * V(i, j - 1) + eta * V(i + 1, j - 1);
                        double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                        M1[GI(i, j)] = (vsp * rho_mix(p, Uc->vof_old[CI(i, j - 1)], Uc->vofS[CI(i, j - 1)]) +
                                        vsn * rho_mix(p, Uc->vof_old[CI(i, j)], Uc->vofS[CI(i, j)])) *
                                       UDX(i);
                    }
                }
                vs = 0.5 * (V(i, j) + V(i, j - 1));
                W2[GI(i, j)] = 0.0;
                M2[GI(i, j)] = 0.0;
                if (j >= NJ) {
                    if (Vc->aN[GI(i, NJ)] >= EPSI) W2[GI(i, j)] = mix_rho(p, in->pvof[CI(i, NJ)], in->pvofS[CI(i, NJ)]);
                    M2[GI(i, j)] = vs * W2[GI(i, j)] * VDX(i);
                } else if (j == 1) {
                    if (Vc->aS[GI(i, j)] >= EPSI) W2[GI(i, j)] = mix_rho(p, in->pvof[CI(i, j)], in->pvofS[CI(i, j)]);
                    M2[GI(i, j)] = vs * W2[GI(i, j)] * VDX(i);
                } else {
                    if (Vc->aS[GI(i, j)] >= EPSI) {
                        double eta = Vc->EtaN[GI(i, j - 1)];
                        vs = (1.0 - eta) * V(i, j - 1) + eta * V(i, j);
                        vs = ((vs - in->vb) * Vc->AlN[GI(i, j - 1)] + in->vb);
                        size_t c = CI(i, j);
                        if (in->pvof[c] > 1.0 - EPSI || in->pvof[c] < EPSI) {
                            W2[GI(i, j)] = rho_mix(p, in->pvof[c], in->pvofS[c]);
                        } else if (in->pvofS[c] < EPSI) {
                            volf = frac(in->pnx[c], in->pny[c], in->pphi[c] + vs * dt / 2.0 * in->pny[c], PDX(i),
                                        fabs(vs * dt));
                            W2[GI(i, j)] = (volf * in->rho_l / in->rho_ref + (1.0 - volf - 0.0) * in->rho_g / in->rho_ref);
                        } else {
                            orc_cellgeocal(in->pnxS[c], in->pnyS[c], in->pphiS[c] + vs * dt / 2.0 * in->pnyS[c],
                                           in->pnx[c], in->pny[c], in->pphi[c] + vs * dt / 2.0 * in->pny[c], PDX(i),
                                           fabs(vs * dt), &vols, &volf);
                            W2[GI(i, j)] = (volf * in->rho_l / in->rho_ref + (1.0 - volf - vols) * in->rho_g / in->rho_ref);
                        }
                        M2[GI(i, j)] = vs * W2[GI(i, j)] * VDX(i);
                    }
                }
            }
        }
    }
}

/* DensityBasedConvectiveFlux, :1751-1869 */
static void density_conv(const PC* p, double* C1, double* C2, const double* W1, const double* W2, int idir)
{
    const orc_pred_in* in = p->in;
    const orc_stag* Uc = &in->U;
    const orc_stag* Vc = &in->V;
    const double vb = in->vb;
    if (idir) {
        for (int i = 1; i <= NI + 1; i++) {
            for (int j = 1; j <= NJ; j++) {
                double uw = 0.5 * (U(i, j) + U(i - 1, j));
                if (i == 1) {
                    C1[GI(i, j)] = sq(uw) * W1[GI(i, j)] * UDY(j);
                } else if (i >= NI) {
                    C1[GI(i, j)] = sq(uw) * W1[GI(i, j)] * UDY(j);
                } else {
                    double eta = Uc->EtaE[GI(i - 1, j)];
                    uw = (1.0 - eta) * U(i - 1, j) + eta * U(i, j);
                    double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                    C1[GI(i, j)] = (uwp * U(i - 1, j) + uwn * U(i, j)) * Uc->AlE[GI(i - 1, j)] * W1[GI(i, j)] * UDY(j);
                }
                uw = 0.5 * (U(i - 1, j + 1) + U(i - 1, j));
                C2[GI(i, j)] = 0.0;
                if (i > NI) {
                    if (Vc->aE[GI(NI, j)] >= EPSI)
                        C2[GI(i, j)] = uw * 0.5 * (V(i - 1, j) + V(i, j)) *
                                       rho_mix(p, Vc->vof_old[CI(NI, j)], Vc->vofS[CI(NI, j)]) * VDY(j);
                } else if (i == 1) {
                    if (Vc->aW[GI(1, j)] >= EPSI)
                        C2[GI(i, j)] = uw * 0.5 * (V(i - 1, j) + V(i, j)) *
                                       rho_mix(p, Vc->vof_old[CI(i, j)], Vc->vofS[CI(i, j)]) * VDY(j);
                } else {
                    double Sy = Uc->SyN[GI(i - 1, j)];
                    double eta = fabs(Vc->fcey[GI(i - 1, j)] + PDY(j) / 2.0 - Uc->ccy[CI(i - 1, j)]) / Sy;
                    if (fabs(eta) >= 1.0) eta = 0.5;
                    uw = (1.0 - eta) * U(i - 1, j) + eta * U(i - 1, j + 1);
                    double uwp = 0.5 * (uw + fabs(uw)), uwn = 0.5 * (uw - fabs(uw));
                    if (Vc->aW[GI(i, j)] >= EPSI)
                        C2[GI(i, j)] = (uwp * V(i - 1, j) * rho_mix(p, Vc->vof_old[CI(i - 1, j)], Vc->vofS[CI(i - 1, j)]) +
                                        uwn * V(i, j) * rho_mix(p, Vc->vof_old[CI(i, j)], Vc->vofS[CI(i, j)])) *
                                       VDY(j);
                }
            }
        }
    } else {
        for (int i = 1; i <= NI; i++) {
            for (int j = 1; j <= NJ + 1; j++) {
                C1[GI(i, j)] = 0.0;
                if (j > NJ) {
                    double vs = 0.5 * (V(i, j - 1) + V(i + 1, j - 1));
                    if (Uc->aN[GI(i, NJ)] >= EPSI)
                        C1[GI(i, j)] = vs * 0.5 * (U(i, j) + U(i, j - 1)) *
                                       rho_mix(p, Uc->vof_old[CI(i, NJ)], Uc->vofS[CI(i, NJ)]) * UDX(i);
                } else if (j == 1) {
                    double vs = 0.5 * (V(i, j - 1) + V(i + 1, j - 1));
                    if (Uc->aS[GI(i, 1)] >= EPSI)
                        C1[GI(i, j)] = vs * 0.5 * (U(i, j) + U(i, j - 1)) *
                                       rho_mix(p, Uc->vof_old[CI(i, j)], Uc->vofS[CI(i, j)]) * UDX(i);
                } else {
                    double Sx = Vc->SxE[GI(i, j - 1)];
                    double eta = fabs(Uc->fcnx[GI(i, j - 1)] + PDX(i) / 2.0 - Vc->ccx[CI(i, j - 1)]) / Sx;
                    if (fabs(eta) >= 1.0) eta = 0.5;
                    double vs = (1.0 - eta) * V(i, j - 1) + eta * V(i + 1, j - 1);
                    double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                    if (Uc->aS[GI(i, j)] >= EPSI)
                        C1[GI(i, j)] = (vsp * U(i, j - 1) * rho_mix(p, Uc->vof_old[CI(i, j - 1)], Uc->vofS[CI(i, j - 1)]) +
                                        vsn * U(i, j) * rho_mix(p, Uc->vof_old[CI(i, j)], Uc->vofS[CI(i, j)])) *
                                       UDX(i);
                }
                if (j >= NJ) {
                    double vs = 0.5 * (V(i, j) + V(i, j - 1));
                    C2[GI(i, j)] = sq(vs) * W2[GI(i, j)] * VDX(i);
                } else if (j == 1) {
                    double vs = 0.5 * (V(i, j) + V(i, j - 1));
                    C2[GI(i, j)] = sq(vs) * W2[GI(i, j)] * VDX(i);
                } else {
                    double eta = Vc->EtaN[GI(i, j - 1)];
                    double vs = (1.0 - eta) * V(i, j - 1) + eta * V(i, j);
                    vs = ((vs - vb) * Vc->AlN[GI(i, j - 1)] + vb);
                    double vsp = 0.5 * (vs + fabs(vs)), vsn = 0.5 * (vs - fabs(vs));
                    C2[GI(i, j)] = (vsp * V(i, j - 1) + vsn * V(i, j)) * W2[GI(i, j)] * VDX(i);
                }
            }
        }
    }
}

/* ---- viscous fluxes --------------------------------------------------------- */

static double nu_face(const PC* p, double vf)
{
    const orc_pred_in* in = p->in;
    return ((1.0 - vf) * in->nu_g / in->nu_ref + vf * in->nu_l / in->nu_ref);
}

/* DiffusiveFlux, :2031-2125.  uvof/vvof: the U/V vof this call reads. */
static void diffusive(const PC* p, const double* uvof, const double* vvof, double* D1, double* D2, double* E1,
                      double* E2, double* X2, int idir)
{
    const orc_pred_in* in = p->in;
    const orc_stag* Uc = &in->U;
    const orc_stag* Vc = &in->V;
    const double tol = 1.0e-24, Rey = in->Re, vb = in->vb;
    size_t ng = (size_t)(NI + 2) * (NJ + 2);
    memset(E1, 0, ng * sizeof(double));
    memset(E2, 0, ng * sizeof(double));
    memset(X2, 0, ng * sizeof(double));
#define EDGEMIX(vf, vs) ((1.0 - (vf) / (1.0 - (vs) + tol)) * in->nu_g / in->nu_ref + (vf) / (1.0 - (vs) + tol) * in->nu_l / in->nu_ref)
    if (idir) {
        for (int j = 1; j <= NJ; j++) {
            D1[GI(1, j)] = EDGEMIX(uvof[CI(1, j)], Uc->vofS[CI(1, j)]) * Uc->aW[GI(1, j)] * UDY(j) / PDX(1) / Rey;
            D2[GI(1, j)] = EDGEMIX(vvof[CI(1, j)], Vc->vofS[CI(1, j)]) * Vc->aW[GI(1, j)] * VDY(j) / PDX(1) / Rey;
            for (int i = 2; i <= NI; i++) {
                double Sx = Uc->SxE[GI(i - 1, j)];
                double Sy = Uc->ccy[CI(i, j)] - Uc->ccy[CI(i - 1, j)];
                double vf = 0.5 * (uvof[CI(i, j)] / (1.0 - Uc->vofS[CI(i, j)] + tol) +
                                   uvof[CI(i - 1, j)] / (1.0 - Uc->vofS[CI(i - 1, j)] + tol));
                D1[GI(i, j)] = nu_face(p, vf) / Rey * Uc->aW[GI(i, j)] * UDY(j) / Sx;
                if (fabs(Sy) >= 1.0e-3 * UDY(j))
                    E1[GI(i, j)] = nu_face(p, vf) / Rey * Uc->aW[GI(i, j)] * Uc->DAlE[GI(i - 1, j)] * UDY(j);
                Sx = Vc->SxE[GI(i - 1, j)];
                Sy = Vc->ccy[CI(i, j)] - Vc->ccy[CI(i - 1, j)];
                vf = 0.5 * (vvof[CI(i, j)] / (1.0 - Vc->vofS[CI(i, j)] + tol) +
                            vvof[CI(i - 1, j)] / (1.0 - Vc->vofS[CI(i - 1, j)] + tol));
                D2[GI(i, j)] = nu_face(p, vf) / Rey * Vc->aW[GI(i, j)] * VDY(j) / Sx;
                if (fabs(Sy) >= 1.0e-3 * VDY(j)) {
                    E2[GI(i, j)] = nu_face(p, vf) / Rey * Vc->aW[GI(i, j)] * Vc->DAlE[GI(i - 1, j)] * VDY(j);
                    X2[GI(i, j)] = vb * E2[GI(i, j)];
                }
            }
            D1[GI(NI + 1, j)] = EDGEMIX(uvof[CI(NI, j)], Uc->vofS[CI(NI, j)]) / Rey * Uc->aE[GI(NI, j)] * UDY(j) / PDX(NI);
            D2[GI(NI + 1, j)] = EDGEMIX(vvof[CI(NI, j)], Vc->vofS[CI(NI, j)]) / Rey * Vc->aE[GI(NI, j)] * VDY(j) / PDX(NI);
        }
    } else {
        for (int i = 1; i <= NI; i++) {
            D1[GI(i, 1)] = EDGEMIX(uvof[CI(i, 1)], Uc->vofS[CI(i, 1)]) / Rey * Uc->aS[GI(i, 1)] * UDX(i) / PDY(1);
            D2[GI(i, 1)] = EDGEMIX(vvof[CI(i, 1)], Vc->vofS[CI(i, 1)]) / Rey * Vc->aS[GI(i, 1)] * VDX(i) / PDY(1);
            for (int j = 2; j <= NJ; j++) {
                double Sx = Uc->ccx[CI(i, j)] - Uc->ccx[CI(i, j - 1)];
                double Sy = Uc->SyN[GI(i, j - 1)];
                double vf = 0.5 * (uvof[CI(i, j - 1)] / (1.0 - Uc->vofS[CI(i, j - 1)] + tol) +
                                   uvof[CI(i, j)] / (1.0 - Uc->vofS[CI(i, j)] + tol));
                D1[GI(i, j)] = nu_face(p, vf) / Rey * Uc->aS[GI(i, j)] * UDX(i) / Sy;
                if (fabs(Sx) >= 1.0e-3 * UDX(i))
                    E1[GI(i, j)] = nu_face(p, vf) / Rey * Uc->aS[GI(i, j)] * Uc->DAlN[GI(i, j - 1)] * UDX(i);
                Sx = Vc->ccx[CI(i, j)] - Vc->ccx[CI(i, j - 1)];
                Sy = Vc->SyN[GI(i, j - 1)];
                vf = 0.5 * (vvof[CI(i, j - 1)] / (1.0 - Vc->vofS[CI(i, j - 1)] + tol) +
                            vvof[CI(i, j)] / (1.0 - Vc->vofS[CI(i, j)] + tol));
                D2[GI(i, j)] = nu_face(p, vf) / Rey * Vc->aS[GI(i, j)] * VDX(i) / Sy;
                if (fabs(Sx) >= 1.0e-3 * VDX(i)) {
                    E2[GI(i, j)] = nu_face(p, vf) / Rey * Vc->aS[GI(i, j)] * Vc->DAlN[GI(i, j - 1)] * VDX(i);
                    X2[GI(i, j)] = vb * E2[GI(i, j)];
                }
            }
            D1[GI(i, NJ + 1)] = EDGEMIX(uvof[CI(i, NJ)], Uc->vofS[CI(i, NJ)]) / Rey * Uc->aN[GI(i, NJ)] * UDX(i) / PDY(NJ);
            D2[GI(i, NJ + 1)] = EDGEMIX(vvof[CI(i, NJ)], Vc->vofS[CI(i, NJ)]) / Rey * Vc->aN[GI(i, NJ)] * VDX(i) / PDY(NJ);
        }
    }
#undef EDGEMIX
}

/* ---- assembly ------------------------------------------------------------------ */

/* Mixing_Procedure, :2127-2197, on the cells i=1..ni-idir, j=1..nj-jdir */
static void mixing(const PC* p, const orc_stag* T, int idir, int jdir, double* fm)
{
    size_t n = (size_t)NI * NJ;
    double* f = (double*)malloc(n * sizeof(double));
    memcpy(f, fm, n * sizeof(double));
    int imax = NI - idir, jmax = NJ - jdir;
    for (int i = 1; i <= imax; i++) {
        for (int j = 1; j <= jmax; j++) {
            size_t c = CI(i, j);
            if (!(T->vofS[c] >= 0.5 + EPSI && T->flag[c])) continue;
            double Fm[3], Be[3], Vt[3];
            int sx = (int)fsign(1.0, T->nxS[c]), sy = (int)fsign(1.0, T->nyS[c]);
            int ii = i + sx, jj = j;
            if (ii <= imax && ii >= 1) {
                Vt[0] = 1.0 - T->vofS[CI(ii, jj)];
                Fm[0] = (1.0 - T->vofS[c]) * Vt[0] * (f[CI(ii, jj)] - f[c]);
                Be[0] = sq(T->nxS[c]) * pow5(Vt[0]);
            } else {
                Fm[0] = 0.0; Be[0] = 0.0; Vt[0] = 1.0;
            }
            ii = i; jj = j + sy;
            if (jj <= jmax && jj >= 1) {
                Vt[1] = 1.0 - T->vofS[CI(ii, jj)];
                Fm[1] = (1.0 - T->vofS[c]) * Vt[1] * (f[CI(ii, jj)] - f[c]);
                Be[1] = sq(T->nyS[c]) * pow5(Vt[1]);
            } else {
                Fm[1] = 0.0; Be[1] = 0.0; Vt[1] = 1.0;
            }
            ii = i + sx; jj = j + sy;
            if (ii <= imax && ii >= 1 && jj <= jmax && jj >= 1) {
                Vt[2] = 1.0 - T->vofS[CI(ii, jj)];
                Fm[2] = (1.0 - T->vofS[c]) * Vt[2] * (f[CI(ii, jj)] - f[c]);
                Be[2] = fabs(T->nxS[c] * T->nyS[c]) * pow5(Vt[2]);
            } else {
                Fm[2] = 0.0; Be[2] = 0.0; Vt[2] = 1.0;
            }
            /* A fully solid target (Vt = 0, so Be = 0) makes the reference's
             * Fmix*Beta/(Beta*(1-vofS)+Vtgt) 0/0, and a cell with no fluid
             * target makes Beta/sum(Beta) 0/0 (NaN in the reference): both
             * take the limit 0 here, the share such a target can hold.  Every
             * other case is the reference's expression unchanged. */
            double temp = (Be[0] + Be[1]) + Be[2];
            for (int k = 0; k < 3; k++) {
                if (temp > 0.0 && Vt[k] > 0.0) {
                    Be[k] = Be[k] / temp;
                    Fm[k] = Fm[k] * Be[k] / (Be[k] * (1.0 - T->vofS[c]) + Vt[k]);
                } else {
                    Fm[k] = 0.0;
                }
            }
            fm[c] = f[c] + 1.0 / (1.0 - T->vofS[c]) * ((Fm[0] + Fm[1]) + Fm[2]);
            ii = i + sx; jj = j;
            if (ii <= imax && ii >= 1 && Vt[0] > 0.0) fm[CI(ii, jj)] = fm[CI(ii, jj)] - 1.0 / Vt[0] * Fm[0];
            ii = i; jj = j + sy;
            if (jj <= jmax && jj >= 1 && Vt[1] > 0.0) fm[CI(ii, jj)] = fm[CI(ii, jj)] - 1.0 / Vt[1] * Fm[1];
            ii = i + sx; jj = j + sy;
            if (ii <= imax && ii >= 1 && jj <= jmax && jj >= 1 && Vt[2] > 0.0)
                fm[CI(ii, jj)] = fm[CI(ii, jj)] - 1.0 / Vt[2] * Fm[2];
        }
    }
    free(f);
}

/* The convective part of the U (comp 0) or V (comp 1) assembly, :175-229 and
 * :320-385: FluxDiv before mixing; updates Flux_n1 on active cells.  C, M,
 * CF hold the component's EW face array, then its NS face array (ng apart):
 * ConFlux, MaFlux, CFlux. */
static void assemble_conv(const PC* p, int comp, const double* C, const double* M, const double* CF,
                          double* fluxn1, double* fd)
{
    const orc_pred_in* in = p->in;
    const orc_stag* T = comp == 0 ? &in->U : &in->V;
    const double* vel = comp == 0 ? in->u : in->v;
    const double rl = in->rho_l, rg = in->rho_g, rr = in->rho_ref;
    const double dt = in->dt;
    size_t ng = (size_t)(NI + 2) * (NJ + 2);
    int imax = NI - (comp == 0), jmax = NJ - (comp == 1);
    for (int i = 1; i <= imax; i++) {
        for (int j = 1; j <= jmax; j++) {
            size_t c = CI(i, j);
            fd[c] = 0.0;
            if (!T->flag[c]) continue;
            double hx = comp == 0 ? UDX(i) : VDX(i), hy = comp == 0 ? UDY(j) : VDY(j);
            double Fe = C[GI(i + 1, j)], Fw = C[GI(i, j)], Fn = C[ng + GI(i, j + 1)], Fs = C[ng + GI(i, j)];
            double vf = T->vof_old[c], vs = T->vofS[c];
            double RoCeO = (1.0 - vf - vs) * rg / rr + vf * rl / rr;
            double RoCeN = RoCeO - dt * (M[GI(i + 1, j)] - M[GI(i, j)] + M[ng + GI(i, j + 1)] - M[ng + GI(i, j)]) / (hx * hy);
            double lo = (1.0 + EPSI - vs) * rg / rr, hi = (1.0 - EPSI - vs) * rl / rr;
            double fluxn0;
            if (RoCeN > lo && RoCeN < hi && RoCeO > lo && RoCeO < hi && vs < EPSI) {
                fluxn0 = (Fe - Fw + Fn - Fs) / RoCeN + vel[GI(i, j)] * (1.0 - RoCeO / RoCeN) / dt * hx * hy;
            } else {
                fluxn0 = (CF[GI(i + 1, j)] - CF[GI(i, j)] + CF[ng + GI(i, j + 1)] - CF[ng + GI(i, j)]) / (1.0 - vs);
            }
            double v0;
            if (!in->first_step)
                v0 = 1.5 * fluxn0 - 0.5 * fluxn1[(size_t)comp * ng + GI(i, j)];
            else
                v0 = fluxn0;
            int closed = 0;
            if (T->aW[GI(i, j)] < EPSI) closed++;
            if (T->aE[GI(i, j)] < EPSI) closed++;
            if (T->aS[GI(i, j)] < EPSI) closed++;
            if (T->aN[GI(i, j)] < EPSI) closed++;
            if (closed >= 3) v0 = 0.0;
            fd[c] = v0 - vel[GI(i, j)] * (1.0 - vs) / (1.0 - vs) * hx * hy / dt;
            fluxn1[(size_t)comp * ng + GI(i, j)] = fluxn0;
        }
    }
}

/* explicit half of the viscous term (+ the moving-wall friction of V), :231-249, :387-420 */
static void assemble_visc(const PC* p, int comp, const double* D1, const double* D2, const double* X1,
                          const double* X2, double* fd)
{
    const orc_pred_in* in = p->in;
    const orc_stag* T = comp == 0 ? &in->U : &in->V;
    const double* w = comp == 0 ? in->u : in->v;
    size_t ng = (size_t)(NI + 2) * (NJ + 2);
    int imax = NI - (comp == 0), jmax = NJ - (comp == 1);
    /* DifFluxEW(.,.,comp) at D1 + comp*ng, DifFluxNS at D2 + comp*ng; ExED likewise */
    const double* DE = D1 + (size_t)comp * ng;
    const double* DN = D2 + (size_t)comp * ng;
    const double* XE = X1 + (size_t)comp * ng;
    const double* XN = X2 + (size_t)comp * ng;
#define W(i, j) w[GI(i, j)]
    for (int i = 1; i <= imax; i++) {
        for (int j = 1; j <= jmax; j++) {
            size_t c = CI(i, j);
            if (!T->flag[c]) continue;
            double ExDe = XE[GI(i + 1, j)], ExDw = XE[GI(i, j)], ExDn = XN[GI(i, j + 1)], ExDs = XN[GI(i, j)];
            double vs = T->vofS[c];
            fd[c] = fd[c] -
                    BETAVIS * ((DE[GI(i + 1, j)] * (W(i + 1, j) - W(i, j)) - DE[GI(i, j)] * (W(i, j) - W(i - 1, j))) -
                               (DN[GI(i, j + 1)] * (W(i, j + 1) - W(i, j)) - DN[GI(i, j)] * (W(i, j) - W(i, j - 1)))) /
                        (1.0 - vs) +
                    (-ExDe + ExDw - ExDn + ExDs) / (1.0 - vs);
            if (comp == 1 && vs > EPSI && vs < 1.0 - EPSI) {
                double vf = T->vof_new[c];
                fd[c] = fd[c] - ((1.0 - vf / (1.0 - vs)) * in->nu_g / in->nu_ref + vf / (1.0 - vs) * in->nu_l / in->nu_ref) /
                                    in->Re * T->WlLh[c] / T->delh[c] * in->vb;
            }
        }
    }
#undef W
}

/* SetMatrix + SetVectors for one component, :622-845: rows in the
 * reference's matr layout (5 planes S,W,diag,E,N), the right-hand side,
 * and Dp on active cells. */
static void momentum_rows(const PC* p, int comp, const double* D1, const double* D2, const double* E1,
                          const double* E2, const double* fd, double* matr, double* rhs, double* Dp)
{
    const orc_pred_in* in = p->in;
    const orc_stag* T = comp == 0 ? &in->U : &in->V;
    const int iu = comp == 0, iv = comp == 1;
    size_t n = (size_t)NI * NJ, ng = (size_t)(NI + 2) * (NJ + 2);
    const double* DE = D1 + (size_t)comp * ng;
    const double* DN = D2 + (size_t)comp * ng;
    const double* EE = E1 + (size_t)comp * ng;
    const double* EN = E2 + (size_t)comp * ng;
    for (size_t k = 0; k < 5 * n; k++) matr[k] = 0.0;
    for (size_t k = 0; k < n; k++) rhs[k] = 0.0;
    double* CWE1 = (double*)calloc((size_t)NJ + 1, sizeof(double));
    double* CWE2 = (double*)calloc((size_t)NJ + 1, sizeof(double));
    double* CSN1 = (double*)calloc((size_t)NI + 1, sizeof(double));
    double* CSN2 = (double*)calloc((size_t)NI + 1, sizeof(double));
    int imax = NI - iu, jmax = NJ - iv;
    for (int i = 1; i <= imax; i++) {
        for (int j = 1; j <= jmax; j++) {
            size_t c = CI(i, j);
            if (!T->flag[c]) continue;
            double hx = iu ? UDX(i) : VDX(i), hy = iu ? UDY(j) : VDY(j);
            double VofAF = 1.0 - T->vofS[c];
            double De = (1.0 - BETAVIS) * DE[GI(i + 1, j)] / VofAF;
            double Dw = (1.0 - BETAVIS) * DE[GI(i, j)] / VofAF;
            double Dn = (1.0 - BETAVIS) * DN[GI(i, j + 1)] / VofAF;
            double Ds = (1.0 - BETAVIS) * DN[GI(i, j)] / VofAF;
            double Fep = EE[GI(i + 1, j)] * (1.0 - T->EtaE[GI(i, j)]) / VofAF;
            double Fem = EE[GI(i + 1, j)] * T->EtaE[GI(i, j)] / VofAF;
            double Fwp = EE[GI(i, j)] * T->EtaE[GI(i - 1, j)] / VofAF;
            double Fwm = EE[GI(i, j)] * (1.0 - T->EtaE[GI(i - 1, j)]) / VofAF;
            double Fnp = EN[GI(i, j + 1)] * (1.0 - T->EtaN[GI(i, j)]) / VofAF;
            double Fnm = EN[GI(i, j + 1)] * T->EtaN[GI(i, j)] / VofAF;
            double Fsp = EN[GI(i, j)] * T->EtaN[GI(i, j - 1)] / VofAF;
            double Fsm = EN[GI(i, j)] * (1.0 - T->EtaN[GI(i, j - 1)]) / VofAF;
            if (iu) {
                if (j == NJ) Dn = 0.0;
                if (i == NI - iu) De = 0.0;
                if (i == 1) Dw = 0.0;
            } else {
                if (j == NJ - iv) Dn = 0.0;
            }
            double aE = De, aW = Dw, aN = Dn, aS = Ds;
            if (i == 1) CWE1[j] = aW;
            if (i == NI - iu) CWE2[j] = aE;
            if (j == 1) CSN1[i] = aS;
            if (j == NJ - iv) CSN2[i] = aN;
            double vf = T->vof_new[c], vs = T->vofS[c];
            double aP = hx * hy / in->dt + aE + aW + aN + aS +
                        ((1.0 - vf / (1.0 - vs)) * in->nu_g / in->nu_ref + vf / (1.0 - vs) * in->nu_l / in->nu_ref) /
                            in->Re * T->WlLh[c] / T->delh[c] / VofAF;
            aP = aP + Fep - Fwp + Fnp - Fsp;
            aE = aE - Fem;
            aW = aW + Fwm;
            aN = aN - Fnm;
            aS = aS + Fsm;
            Dp[GI(i, j)] = hx * hy / (aP - aE - aW - aN - aS);
            if (j > 1 && T->flag[CI(i, j - 1)]) matr[0 * n + c] = -aS;
            if (i > 1 && T->flag[CI(i - 1, j)]) matr[1 * n + c] = -aW;
            matr[2 * n + c] = aP;
            if (i < NI && T->flag[CI(i + 1, j)]) matr[3 * n + c] = -aE;
            if (j < NJ && T->flag[CI(i, j + 1)]) matr[4 * n + c] = -aN;
        }
    }
    for (int i = 1; i <= imax; i++) {
        for (int j = 1; j <= jmax; j++) {
            size_t c = CI(i, j);
            if (!T->flag[c]) continue;
            double hx = iu ? UDX(i) : VDX(i), hy = iu ? UDY(j) : VDY(j);
            double r = 0.0;
            r = r - fd[c];
            if (iv) r = r - 1.0 / (in->Fr * in->Fr) * hx * hy;
            const double* uu = in->u;
            const double* vv = in->v;
            if (i == 1) r = r + CWE1[j] * ((double)iu * uu[GI(i - 1, j)] + (double)iv * vv[GI(i - 1, j)]);
            if (i == NI - iu) r = r + CWE2[j] * ((double)iu * uu[GI(i + 1, j)] + (double)iv * vv[GI(i + 1, j)]);
            if (j == 1) r = r + CSN1[i] * ((double)iu * uu[GI(i, j - 1)] + (double)iv * vv[GI(i, j - 1)]);
            if (j == NJ - iv) r = r + CSN2[i] * ((double)iu * uu[GI(i, j + 1)] + (double)iv * vv[GI(i, j + 1)]);
            rhs[c] = r;
        }
    }
    free(CWE1); free(CWE2); free(CSN1); free(CSN2);
}

/* Predictor_UV, PredictorUV_mod.f90:28-497 up to the two linear systems:
 * returns the U and V momentum rows (matr layout, 5 planes), right-hand
 * sides, PU%Dp / PV%Dp on the active cells, and the updated Flux_n1
 * (ghosted, 2 planes, in/out).  The systems are solved by the caller. */
void orc_predictor_assemble(const orc_pred_in* in, double* fluxn1, double* matr_u, double* rhs_u, double* matr_v,
                            double* rhs_v, double* DpU, double* DpV)
{
    PC pc = {in, in->ni, in->nj}, *p = &pc;
    size_t ng = (size_t)(NI + 2) * (NJ + 2), n = (size_t)NI * NJ;
    double* CF1 = (double*)calloc(2 * ng, sizeof(double)); /* CFluxEW, comps 1,2 */
    double* CF2 = (double*)calloc(2 * ng, sizeof(double)); /* CFluxNS */
    double* M1 = (double*)calloc(2 * ng, sizeof(double));  /* MaFluxEW */
    double* M2 = (double*)calloc(2 * ng, sizeof(double));  /* MaFluxNS */
    double* W1 = (double*)calloc(2 * ng, sizeof(double));  /* VofFluxEW */
    double* W2 = (double*)calloc(2 * ng, sizeof(double));  /* VofFluxNS */
    double* K1 = (double*)calloc(2 * ng, sizeof(double));  /* ConFluxEW */
    double* K2 = (double*)calloc(2 * ng, sizeof(double));  /* ConFluxNS */
    double* D1 = (double*)calloc(2 * ng, sizeof(double));  /* DifFluxEW */
    double* D2 = (double*)calloc(2 * ng, sizeof(double));  /* DifFluxNS */
    double* E1 = (double*)calloc(2 * ng, sizeof(double));  /* EDFluxEW */
    double* E2 = (double*)calloc(2 * ng, sizeof(double));  /* EDFluxNS */
    double* X1 = (double*)calloc(2 * ng, sizeof(double));  /* ExEDFluxEW */
    double* X2 = (double*)calloc(2 * ng, sizeof(double));  /* ExEDFluxNS */
    double* fdu = (double*)calloc(n, sizeof(double));
    double* fdv = (double*)calloc(n, sizeof(double));
    /* Step 1: convective fluxes */
    if (in->first_step) {
        modified_conv(p, CF1, CF1 + ng, 1);
        modified_conv(p, CF2, CF2 + ng, 0);
    } else {
        high_order_x(p, CF1, CF1 + ng);
        high_order_y(p, CF2, CF2 + ng);
    }
    face_density(p, M1, M1 + ng, W1, W1 + ng, 1);
    face_density(p, M2, M2 + ng, W2, W2 + ng, 0);
    density_conv(p, K1, K1 + ng, W1, W1 + ng, 1);
    density_conv(p, K2, K2 + ng, W2, W2 + ng, 0);
    /* Step 2: diffusive fluxes at time n */
    diffusive(p, in->U.vof_old, in->V.vof_old, D1, D1 + ng, E1, E1 + ng, X1 + ng, 1);
    diffusive(p, in->U.vof_old, in->V.vof_old, D2, D2 + ng, E2, E2 + ng, X2 + ng, 0);
    /* ConFlux(EW, comp) lives at K1 + comp*ng; assemble_conv reads EW at
     * C[..] and NS at C[ng + ..] of a per-component pair */
    {
        double* CU = (double*)malloc(2 * ng * sizeof(double));
        double* MU = (double*)malloc(2 * ng * sizeof(double));
        double* FU = (double*)malloc(2 * ng * sizeof(double));
        for (int comp = 0; comp < 2; comp++) {
            memcpy(CU, K1 + comp * ng, ng * sizeof(double));
            memcpy(CU + ng, K2 + comp * ng, ng * sizeof(double));
            memcpy(MU, M1 + comp * ng, ng * sizeof(double));
            memcpy(MU + ng, M2 + comp * ng, ng * sizeof(double));
            memcpy(FU, CF1 + comp * ng, ng * sizeof(double));
            memcpy(FU + ng, CF2 + comp * ng, ng * sizeof(double));
            double* fd = comp == 0 ? fdu : fdv;
            assemble_conv(p, comp, CU, MU, FU, fluxn1, fd);
            mixing(p, comp == 0 ? &in->U : &in->V, comp == 0, comp == 1, fd);
            assemble_visc(p, comp, D1, D2, X1, X2, fd);
        }
        free(CU); free(MU); free(FU);
    }
    /* implicit part: diffusive fluxes with the new U/V vof */
    diffusive(p, in->U.vof_new, in->V.vof_new, D1, D1 + ng, E1, E1 + ng, X1 + ng, 1);
    diffusive(p, in->U.vof_new, in->V.vof_new, D2, D2 + ng, E2, E2 + ng, X1 + ng, 0);
    momentum_rows(p, 0, D1, D2, E1, E2, fdu, matr_u, rhs_u, DpU);
    momentum_rows(p, 1, D1, D2, E1, E2, fdv, matr_v, rhs_v, DpV);
    free(CF1); free(CF2); free(M1); free(M2); free(W1); free(W2); free(K1); free(K2);
    free(D1); free(D2); free(E1); free(E2); free(X1); free(X2); free(fdu); free(fdv);
}

/* The ghost copies of PU%Dp / PV%Dp after the solves (:530-540) and the
 * open-top copy of PredictorVelocityBoundaryCondition (:2199-2213). */
void orc_predictor_finish(int ni, int nj, double* DpU, double* DpV, double* vpred)
{
    PC pc = {NULL, ni, nj}, *p = &pc;
#define DU(i, j) DpU[GI(i, j)]
#define DV(i, j) DpV[GI(i, j)]
    for (int j = 0; j <= NJ + 1; j++) {
        DU(0, j) = DU(1, j);
        DU(NI, j) = DU(NI - 1, j);
        DU(NI + 1, j) = DU(NI, j);
    }
    for (int i = 0; i <= NI + 1; i++) {
        DU(i, 0) = DU(i, 1);
        DU(i, NJ + 1) = DU(i, NJ);
    }
    for (int j = 0; j <= NJ + 1; j++) {
        DV(0, j) = DV(1, j);
        DV(NI + 1, j) = DV(NI, j);
    }
    for (int i = 0; i <= NI + 1; i++) {
        DV(i, 0) = DV(i, 1);
        DV(i, NJ) = DV(i, NJ - 1);
        DV(i, NJ + 1) = DV(i, NJ);
    }
    for (int i = 1; i <= NI; i++) vpred[GI(i, NJ)] = vpred[GI(i, NJ - 1)];
#undef DU
#undef DV
}

/* ---- test hooks: single pieces of the predictor for the closed-form checks
 * (no counterpart in the reference) ---------------------------------------- */
double orc_muscl_limiter(double r, double tolim) { return muscl(r, tolim); }

/* HighOrderConvectiveFluxForXDir / YDir (itt > 1): CFluxEW (2 ghosted
 * planes, comps 1,2) and CFluxNS */
void orc_conv_fluxes(const orc_pred_in* in, double* CF1, double* CF2)
{
    PC pc = {in, in->ni, in->nj}, *p = &pc;
    size_t ng = (size_t)(NI + 2) * (NJ + 2);
    high_order_x(p, CF1, CF1 + ng);
    high_order_y(p, CF2, CF2 + ng);
}

/* Mixing_Procedure of the U (comp 0) or V (comp 1) cells on a cell array */
void orc_mixing(const orc_pred_in* in, int comp, double* flux)
{
    PC pc = {in, in->ni, in->nj}, *p = &pc;
    mixing(p, comp == 0 ? &in->U : &in->V, comp == 0, comp == 1, flux);
}
