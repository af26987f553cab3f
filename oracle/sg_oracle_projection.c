/*
 * CPU FP64 oracle: projection step (coefficients, matrix, RHS, residual,
 * pressure scatter/ghosts, velocity correction).
 *
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP path, never part of it.
 *
 * Loop orders and floating-point evaluation orders follow the reference
 * line by line (Fortran evaluates a*b/c*d left to right, as C does); the
 * library is built with -ffp-contract=off so that no FMA changes a rounding.
 */
#include "sg_oracle.h"
#include <math.h>
#include <string.h>

#define EPSI 1.0e-3   /* StateVariables.f90:28 */
#define TOLB 1.0e-24  /* ProjectionP_Mod.f90:404, ComputePUV_Mod.f90:46 */

/* cell (1-based) and ghosted (0-based origin) indexing */
#define C(a, i, j) (a)[((i) - 1) + ((j) - 1) * (size_t)ni]
#define G(a, i, j) (a)[(i) + (j) * (size_t)(ni + 2)]
#define PC(i, j, k) PoCoef[((i) - 1) + ((j) - 1) * (size_t)ni + ((k) - 1) * (size_t)ni * nj]
#define MT(i, j, k) matr[((i) - 1) + ((j) - 1) * (size_t)ni + ((k) - 1) * (size_t)ni * nj]

/* Fortran DSIGN(1.d0, x): +1 for x >= 0 (and +0), -1 for x < 0. */
static double fsign1(double x) { return (x < 0.0) ? -1.0 : 1.0; }

/* ProjectionP_Mod.f90:409-410 (wet) -- the dry test of :493-494 is its
 * exact complement for finite inputs. */
int orc_is_wet(double vof, double phi, double vofS) {
    return (vof >= 0.5 && vofS < EPSI) || (phi < 0.0 && vofS >= EPSI);
}
static int is_dry(double vof, double phi, double vofS) {
    return (vof < 0.5 && vofS < EPSI) || (phi >= 0.0 && vofS >= EPSI);
}

/* ProjectionP_Mod.f90:395-580.  BetaW and BetaD are carried across cells
 * exactly like the reference's local variables, so the dry/dry branch of
 * :504-508 (which assigns BetaW but divides by the previous BetaD) is
 * reproduced with the same stale value. */
void orc_poisson_coef(int ni, int nj, const double* dx, const double* dy,
                      const double* udx, const double* vdy,
                      const double* vof, const double* phi, const double* vofS,
                      const double* DpU, const double* DpV,
                      double rho_l, double rho_g, double rho_ref,
                      double* PoCoef) {
    const double tol = TOLB;
    const double BetaP = 1.0 / (rho_l / rho_ref);
    const double BetaM = 1.0 / (rho_g / rho_ref);
    double BetaW = 0.0, BetaD = 0.0, Lamda;
    memset(PoCoef, 0, sizeof(double) * (size_t)ni * nj * 4);
#define WET(i, j) orc_is_wet(C(vof, i, j), C(phi, i, j), C(vofS, i, j))
#define DRY(i, j) is_dry(C(vof, i, j), C(phi, i, j), C(vofS, i, j))
#define LAM(i2, j2) (fabs(C(phi, i, j)) / (fabs(C(phi, i, j)) + fabs(C(phi, i2, j2)) + tol))
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            const double dxi = dx[i - 1], dyj = dy[j - 1];
            if (WET(i, j)) {
                /* South (:412-431) */
                if (j > 1) {
                    if (DRY(i, j - 1)) {
                        Lamda = LAM(i, j - 1);
                        BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                        PC(i, j, 1) = G(DpV, i, j - 1) * dxi / vdy[j - 2] * BetaP * BetaM / BetaW;
                    } else if (WET(i, j - 1)) {
                        BetaW = BetaM;
                        PC(i, j, 1) = G(DpV, i, j - 1) * dxi / vdy[j - 2] * BetaP * BetaM / BetaW;
                    }
                } else {
                    BetaW = BetaM;
                    PC(i, j, 1) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaW;
                }
                /* West (:433-451) */
                if (i > 1) {
                    if (DRY(i - 1, j)) {
                        Lamda = LAM(i - 1, j);
                        BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                        PC(i, j, 2) = G(DpU, i - 1, j) * dyj / udx[i - 2] * BetaP * BetaM / BetaW;
                    } else if (WET(i - 1, j)) {
                        BetaW = BetaM;
                        PC(i, j, 2) = G(DpU, i - 1, j) * dyj / udx[i - 2] * BetaP * BetaM / BetaW;
                    }
                } else {
                    BetaW = BetaM;
                    PC(i, j, 2) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaW;
                }
                /* East (:454-472) */
                if (i < ni) {
                    if (DRY(i + 1, j)) {
                        Lamda = LAM(i + 1, j);
                        BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                        PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaW;
                    } else if (WET(i + 1, j)) {
                        BetaW = BetaM;
                        PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaW;
                    }
                } else {
                    BetaW = BetaM;
                    PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaW;
                }
                /* North (:474-492) */
                if (j < nj) {
                    if (DRY(i, j + 1)) {
                        Lamda = LAM(i, j + 1);
                        BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                        PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaW;
                    } else if (WET(i, j + 1)) {
                        BetaW = BetaM;
                        PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaW;
                    }
                } else {
                    BetaW = BetaM;
                    PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaW;
                }
            } else if (DRY(i, j)) {
                /* South (:495-514): the dry/dry branch sets BetaW but uses
                 * BetaD, whatever the previous assignment left in it. */
                if (j > 1) {
                    if (WET(i, j - 1)) {
                        Lamda = LAM(i, j - 1);
                        BetaD = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                        PC(i, j, 1) = G(DpV, i, j - 1) * dxi / vdy[j - 2] * BetaP * BetaM / BetaD;
                    } else if (DRY(i, j - 1)) {
                        BetaW = BetaP;
                        PC(i, j, 1) = G(DpV, i, j - 1) * dxi / vdy[j - 2] * BetaP * BetaM / BetaD;
                    }
                } else {
                    BetaD = BetaP;
                    PC(i, j, 1) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaD;
                }
                /* West (:515-534) */
                if (i > 1) {
                    if (WET(i - 1, j)) {
                        Lamda = LAM(i - 1, j);
                        BetaD = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                        PC(i, j, 2) = G(DpU, i - 1, j) * dyj / udx[i - 2] * BetaP * BetaM / BetaD;
                    } else if (DRY(i - 1, j)) {
                        BetaD = BetaP;
                        PC(i, j, 2) = G(DpU, i - 1, j) * dyj / udx[i - 2] * BetaP * BetaM / BetaD;
                    }
                } else {
                    BetaD = BetaP;
                    PC(i, j, 2) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaD;
                }
                /* East (:537-555) */
                if (i < ni) {
                    if (WET(i + 1, j)) {
                        Lamda = LAM(i + 1, j);
                        BetaD = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                        PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaD;
                    } else if (DRY(i + 1, j)) {
                        BetaD = BetaP;
                        PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaD;
                    }
                } else {
                    BetaD = BetaP;
                    PC(i, j, 3) = G(DpU, i, j) * dyj / udx[i - 1] * BetaP * BetaM / BetaD;
                }
                /* North (:557-576) */
                if (j < nj) {
                    if (WET(i, j + 1)) {
                        Lamda = LAM(i, j + 1);
                        BetaD = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                        PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaD;
                    } else if (DRY(i, j + 1)) {
                        BetaD = BetaP;
                        PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaD;
                    }
                } else {
                    BetaD = BetaP;
                    PC(i, j, 4) = G(DpV, i, j) * dxi / vdy[j - 1] * BetaP * BetaM / BetaD;
                }
            }
        }
    }
    (void)BetaW;
#undef WET
#undef DRY
#undef LAM
}

/* ProjectionP_Mod.f90:167-266.  Posnu /= -1  <=>  vofS < 1-epsi
 * (Cutcell_Mod.f90:349). */
void orc_poisson_matrix(int ni, int nj, const double* PoCoef,
                        const double* aE, const double* aW,
                        const double* aN, const double* aS,
                        const double* vofS, int bcW, int bcE, int bcS, int bcN,
                        double* matr) {
    const double tol = TOLB;
    memset(matr, 0, sizeof(double) * (size_t)ni * nj * 5);
#define FLUID(i, j) (C(vofS, i, j) < 1.0 - EPSI)
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            if (!FLUID(i, j)) continue;
            if (j > 1 && FLUID(i, j - 1)) MT(i, j, 1) = -PC(i, j, 1) * G(aS, i, j);
            if (i > 1 && FLUID(i - 1, j)) MT(i, j, 2) = -PC(i, j, 2) * G(aW, i, j);
            double d = PC(i, j, 1) * G(aS, i, j) + PC(i, j, 2) * G(aW, i, j) +
                       PC(i, j, 3) * G(aE, i, j) + PC(i, j, 4) * G(aN, i, j);
            if (bcS == 1 && j == 1) d = d - PC(i, j, 1) * G(aS, i, j);
            if (bcW == 1 && i == 1) d = d - PC(i, j, 2) * G(aW, i, j);
            if (bcE == 1 && i == ni) d = d - PC(i, j, 3) * G(aE, i, j);
            if (bcN == 1 && j == nj) d = d - PC(i, j, 4) * G(aN, i, j);
            d = d + fsign1(d) * tol;
            MT(i, j, 3) = d;
            if (i < ni && FLUID(i + 1, j)) MT(i, j, 4) = -PC(i, j, 3) * G(aE, i, j);
            if (j < nj && FLUID(i, j + 1)) MT(i, j, 5) = -PC(i, j, 4) * G(aN, i, j);
        }
    }
#undef FLUID
}

/* ProjectionP_Mod.f90:306-338 */
void orc_poisson_rhs(int ni, int nj, const double* dx, const double* dy,
                     const double* u, const double* v,
                     const double* aE, const double* aW,
                     const double* aN, const double* aS,
                     const double* nyS, const double* WlLh, const double* vofS,
                     double vb, double* rhs) {
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            if (C(vofS, i, j) < 1.0 - EPSI) {
                const double dxi = dx[i - 1], dyj = dy[j - 1];
                C(rhs, i, j) = -(G(u, i, j) * G(aE, i, j) - G(u, i - 1, j) * G(aW, i, j)) * dyj -
                               (G(v, i, j) * G(aN, i, j) - G(v, i, j - 1) * G(aS, i, j)) * dxi +
                               vb * C(nyS, i, j) * C(WlLh, i, j);
            } else {
                C(rhs, i, j) = 0.0;
            }
        }
    }
}

/* ProjectionP_Mod.f90:70-85 */
void orc_matr_apply(int ni, int nj, const double* matr, const double* x, double* y) {
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            double resi = 0.0;
            if (j > 1) resi = resi + MT(i, j, 1) * C(x, i, j - 1);
            if (i > 1) resi = resi + MT(i, j, 2) * C(x, i - 1, j);
            resi = resi + MT(i, j, 3) * C(x, i, j);
            if (i < ni) resi = resi + MT(i, j, 4) * C(x, i + 1, j);
            if (j < nj) resi = resi + MT(i, j, 5) * C(x, i, j + 1);
            C(y, i, j) = resi;
        }
    }
}

double orc_max_row_residual(int ni, int nj, const double* matr, const double* x,
                            const double* b, const double* vofS) {
    double m = 0.0;
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            if (!(C(vofS, i, j) < 1.0 - EPSI)) continue;
            double resi = 0.0;
            if (j > 1) resi = resi + MT(i, j, 1) * C(x, i, j - 1);
            if (i > 1) resi = resi + MT(i, j, 2) * C(x, i - 1, j);
            resi = resi + MT(i, j, 3) * C(x, i, j);
            if (i < ni) resi = resi + MT(i, j, 4) * C(x, i + 1, j);
            if (j < nj) resi = resi + MT(i, j, 5) * C(x, i, j + 1);
            double d = fabs(resi - C(b, i, j));
            if (d > m) m = d;
        }
    }
    return m;
}

/* ProjectionP_Mod.f90:377-391 then :582-595.  Corners keep the zero that
 * UpdatePUV initialises Pp with (ComputePUV_Mod.f90:50). */
void orc_pressure_scatter_bc(int ni, int nj, const double* x, const double* vofS,
                             int bcW, int bcE, int bcS, int bcN, double* Pp) {
    memset(Pp, 0, sizeof(double) * (size_t)(ni + 2) * (nj + 2));
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            G(Pp, i, j) = (C(vofS, i, j) < 1.0 - EPSI) ? C(x, i, j) : 0.0;
    for (int i = 1; i <= ni; ++i) {
        G(Pp, i, 0) = (double)bcS * G(Pp, i, 1);
        G(Pp, i, nj + 1) = (double)bcN * G(Pp, i, nj);
    }
    for (int j = 1; j <= nj; ++j) {
        G(Pp, 0, j) = (double)bcW * G(Pp, 1, j);
        G(Pp, ni + 1, j) = (double)bcE * G(Pp, ni, j);
    }
}

/* ComputePUV_Mod.f90:59-217 and VariablesInternalCellCondition :227-245. */
void orc_velocity_correct(int ni, int nj, const double* dx, const double* dy,
                          const double* udx, const double* vdy,
                          const double* vof, const double* phi,
                          const double* vofS, const double* nyS,
                          const double* WlLh,
                          const double* aE, const double* aW,
                          const double* aN, const double* aS,
                          const int* uflag, const int* vflag,
                          const double* vofSU, const double* vofSV,
                          const double* DpU, const double* DpV,
                          const double* upred, const double* vpred,
                          const double* Pp, double vb,
                          double rho_l, double rho_g, double rho_ref,
                          double* u, double* v, double* p, double* mres) {
    const double tol = TOLB;
    const double BetaP = 1.0 / (rho_l / rho_ref);
    const double BetaM = 1.0 / (rho_g / rho_ref);
    double BetaW, Lamda, grad;
#define WET(i, j) orc_is_wet(C(vof, i, j), C(phi, i, j), C(vofS, i, j))
#define DRY(i, j) is_dry(C(vof, i, j), C(phi, i, j), C(vofS, i, j))
    /* u faces :59-100 */
    for (int i = 1; i <= ni - 1; ++i) {
        for (int j = 1; j <= nj; ++j) {
            if (!C(uflag, i, j)) continue;
            int set = 0;
            if (WET(i, j)) {
                if (DRY(i + 1, j)) {
                    Lamda = fabs(C(phi, i, j)) / (fabs(C(phi, i, j)) + fabs(C(phi, i + 1, j)) + tol);
                    BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                    set = 1;
                } else if (WET(i + 1, j)) {
                    BetaW = BetaM;
                    set = 1;
                }
            } else if (DRY(i, j)) {
                if (WET(i + 1, j)) {
                    Lamda = fabs(C(phi, i, j)) / (fabs(C(phi, i, j)) + fabs(C(phi, i + 1, j)) + tol);
                    BetaW = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                    set = 1;
                } else if (DRY(i + 1, j)) {
                    BetaW = BetaP;
                    set = 1;
                }
            }
            if (set) {
                grad = (G(Pp, i + 1, j) - G(Pp, i, j)) * BetaM * BetaP / BetaW * G(DpU, i, j) / udx[i - 1];
                G(u, i, j) = G(upred, i, j) - grad;
            }
        }
    }
    /* v faces :111-152 */
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj - 1; ++j) {
            if (!C(vflag, i, j)) continue;
            int set = 0;
            if (WET(i, j)) {
                if (DRY(i, j + 1)) {
                    Lamda = fabs(C(phi, i, j)) / (fabs(C(phi, i, j)) + fabs(C(phi, i, j + 1)) + tol);
                    BetaW = Lamda * BetaM + (1.0 - Lamda) * BetaP;
                    set = 1;
                } else if (WET(i, j + 1)) {
                    BetaW = BetaM;
                    set = 1;
                }
            } else if (DRY(i, j)) {
                if (WET(i, j + 1)) {
                    Lamda = fabs(C(phi, i, j)) / (fabs(C(phi, i, j)) + fabs(C(phi, i, j + 1)) + tol);
                    BetaW = Lamda * BetaP + (1.0 - Lamda) * BetaM;
                    set = 1;
                } else if (DRY(i, j + 1)) {
                    BetaW = BetaP;
                    set = 1;
                }
            }
            if (set) {
                grad = (G(Pp, i, j + 1) - G(Pp, i, j)) * BetaM * BetaP / BetaW * G(DpV, i, j) / vdy[j - 1];
                G(v, i, j) = G(vpred, i, j) - grad;
            }
        }
    }
#undef WET
#undef DRY
    /* p <- Pp on fluid cells :163-169 */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            if (C(vofS, i, j) < 1.0 - EPSI) G(p, i, j) = G(Pp, i, j);
    /* top-row v from continuity :177-182 */
    for (int i = 1; i <= ni; ++i) {
        const int j = nj;
        G(v, i, j) = ((-G(u, i, j) * G(aE, i, j) + G(u, i - 1, j) * G(aW, i, j)) * dy[j - 1] +
                      G(v, i, j - 1) * G(aS, i, j) * dx[i - 1]) /
                     dx[i - 1] / G(aN, i, j);
    }
    /* mass residual :184-217 */
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            C(mres, i, j) = (G(u, i, j) * G(aE, i, j) - G(u, i - 1, j) * G(aW, i, j)) * dy[j - 1] +
                            (G(v, i, j) * G(aN, i, j) - G(v, i, j - 1) * G(aS, i, j)) * dx[i - 1] -
                            vb * C(nyS, i, j) * C(WlLh, i, j);
            if (C(vofS, i, j) > 1.0 - EPSI) C(mres, i, j) = 0.0;
        }
    }
    /* VariablesInternalCellCondition :232-244 */
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj; ++j) {
            if (C(vofS, i, j) > 1.0 - EPSI) G(p, i, j) = 0.0;
            if (C(vofSU, i, j) >= 1.0 - EPSI) G(u, i, j) = 0.0;
            if (C(vofSV, i, j) >= 1.0 - EPSI) G(v, i, j) = 0.0;
        }
    }
}
