/*
 * CPU FP64 oracle of the cut-cell geometry of the P, U and V cells.
 * TEST INFRASTRUCTURE ONLY (see sg_oracle.h).
 *
 * Restates, in the reference's loop and expression order:
 *   orc_edge_geo_cal        Edge_Geo_Cal                Cutcell_Mod.f90:302-337
 *   orc_cell_geo_cal        Cell_Geo_Cal                Cutcell_Mod.f90:34-300
 *   orc_define_moex         DefineMomentumExchangeCell  Cutcell_Mod.f90:360-397
 *   orc_number_external     NumberExternalCell          Cutcell_Mod.f90:340-358
 *   orc_new_cell_face       NewCellFace                 Cutcell_Mod.f90:399-549
 *     east_interpolation    EastFaceInterpolationInf    Cutcell_Mod.f90:551-616
 *     north_interpolation   NorthFaceInterpolationInf   Cutcell_Mod.f90:618-700
 *   orc_solid_fraction      frac + clipping             Clsvof_Mod.f90:273-277
 *   orc_cut_cell_geometry   Grid_Preprocess (:18-32) followed by NewCellFace,
 *                           the order of Solver_Mod.f90:303-304.
 *
 * Where the reference reads storage that no statement of these routines (or
 * of a prior call) defines, the oracle fixes a value and says so:
 *   - aperture ghosts (0 and ni+1 / nj+1) are 0 unless a closing cell zeroes
 *     them (the same 0);
 *   - Posnu of the U cells at i = Isize and of the V cells at j = Jsize
 *     (NumberExternalCell stops one short, :347-348) keeps the 0 of a fresh
 *     allocation: those cells count as numbered (/= -1) wherever Posnu is
 *     tested, as in the transport's face masks (Clsvof_Mod.f90:778-797),
 *     while the momentum systems have no row there;
 *   - the face-factor ghosts neither interpolation routine writes (rows 0
 *     and nj+1 and column ni+1 of the east factors, columns 0 and ni+1 and
 *     row nj+1 of the north factors) take the nearest entry that is
 *     written, rows first;
 *   - AlE(0,j) = AlN(1,j) (:606) reads the north factor before
 *     NorthFaceInterpolationInf of the same call rewrites it: the caller
 *     passes AlN holding the previous call's values (in/out);
 *   - MsCe (CellLinking, :702-725) is not formed: nothing on the path reads
 *     it.
 * Arrays: cell arrays (i-1)+(j-1)*ni; ghosted arrays i+j*(ni+2).
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "sg_oracle.h"

#define EPSI 1.0e-3      /* StateVariables.f90:28 */
#define EPSIF 1.0e-2     /* StateVariables.f90:28 */
#define ROUNDOFF 1.0e-24 /* Cutcell_Mod.f90:45 */

#define CL(a, i, j) (a)[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)ni]
#define GH(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(ni + 2)]

/* Cutcell_Mod.f90:302-337.  Returns the SOLID fraction of the edge
 * Pt1-Pt2 and the centre of its solid part (the midpoint when the edge is
 * all fluid or all solid). */
void orc_edge_geo_cal(double p1x, double p1y, double p2x, double p2y, double nxx, double nyy, double diss,
                      double* area, double* fx, double* fy) {
    const double l1 = nxx * p1x + nyy * p1y + diss;
    const double l2 = nxx * p2x + nyy * p2y + diss;
    const double epsil = 1.0e-14;
    double a = 0.0;
    *fx = 0.5 * (p1x + p2x);
    *fy = 0.5 * (p1y + p2y);
    if (l1 >= epsil && l2 >= epsil) {
        a = 0.0;
    } else if (l1 < epsil && l2 < epsil) {
        a = 1.0;
    } else if (l1 >= epsil && l2 < epsil) {
        a = fabs(l2) / (fabs(l1) + fabs(l2));
        if (fabs(p1x - p2x) <= 1.0e-10)
            *fy = p2y - 0.5 * a * (p2y - p1y);
        else
            *fx = p2x - 0.5 * a * (p2x - p1x);
    } else if (l1 < epsil && l2 >= epsil) {
        a = fabs(l1) / (fabs(l1) + fabs(l2));
        if (fabs(p1x - p2x) <= 1.0e-10)
            *fy = p1y - 0.5 * a * (p1y - p2y);
        else
            *fx = p1x - 0.5 * a * (p1x - p2x);
    }
    *area = a;
}

/* solid volume fraction of each cell from its solid line (frac, then the
 * clipping of Clsvof_Mod.f90:273-277) */
void orc_solid_fraction(int ni, int nj, const double* dx, const double* dy, const double* phiS, const double* nxS,
                        const double* nyS, double* vofS) {
    const double vofeps = 1.0e-14; /* Clsvof_Mod.f90:10 */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            double v = orc_frac(CL(nxS, i, j), CL(nyS, i, j), CL(phiS, i, j), dx[i - 1], dy[j - 1]);
            if (v < vofeps) v = 0.0;
            if (v >= 1.0 - vofeps) v = 1.0;
            CL(vofS, i, j) = v;
        }
}

static void close_cell(int ni, orc_ccell* C, int i, int j) {
    CL(C->vofS, i, j) = 1.0;
    GH(C->aW, i, j) = 0.0;
    GH(C->aE, i, j) = 0.0;
    GH(C->aN, i, j) = 0.0;
    GH(C->aS, i, j) = 0.0;
    GH(C->aE, i - 1, j) = 0.0;
    GH(C->aN, i, j - 1) = 0.0;
    GH(C->aW, i + 1, j) = 0.0;
    GH(C->aS, i, j + 1) = 0.0;
}

/* Cutcell_Mod.f90:34-300 for one grid.  Inputs: vofS, phiS, nxS, nyS
 * (cell arrays).  Outputs: apertures, FCE/FCN, WlLh (and delh of full solid
 * cells when C->delh is set), vofS closed; Cell_Cent zeroed when set.
 * C->moex == NULL marks the P cells: the closing rule (MsCe not allocated,
 * :229-232) and the corner patch (MoExCell not allocated, :259) apply. */
void orc_cell_geo_cal(const orc_cgrid* G, orc_ccell* C, int icorprob) {
    const int ni = G->ni, nj = G->nj;
    const int p_cell = C->moex == NULL;
    const size_t nc = (size_t)ni * (size_t)nj;
    double* fcwx = (double*)malloc(nc * sizeof(double));
    double* fcwy = (double*)malloc(nc * sizeof(double));
    double* fcsx = (double*)malloc(nc * sizeof(double));
    double* fcsy = (double*)malloc(nc * sizeof(double));
#define FW(a, i, j) (a)[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)ni]
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            const double dx = G->dx[i - 1], dy = G->dy[j - 1];
            if (C->ccx) {
                CL(C->ccx, i, j) = 0.0;
                CL(C->ccy, i, j) = 0.0;
            }
            double ptx[2][2], pty[2][2];
            for (int ii = 0; ii <= 1; ++ii)
                for (int jj = 0; jj <= 1; ++jj) {
                    ptx[ii][jj] = -(dx * (0.5 - (double)ii));
                    pty[ii][jj] = -(dy * (0.5 - (double)jj));
                }
            const double nx = CL(C->nxS, i, j), ny = CL(C->nyS, i, j), ph = CL(C->phiS, i, j);
            double a, fx, fy;
            orc_edge_geo_cal(ptx[0][0], pty[0][0], ptx[0][1], pty[0][1], nx, ny, ph, &a, &fx, &fy);
            GH(C->aW, i, j) = a;
            FW(fcwx, i, j) = fx;
            FW(fcwy, i, j) = fy;
            orc_edge_geo_cal(ptx[0][1], pty[0][1], ptx[1][1], pty[1][1], nx, ny, ph, &a, &fx, &fy);
            GH(C->aN, i, j) = a;
            GH(C->fcnx, i, j) = fx;
            GH(C->fcny, i, j) = fy;
            orc_edge_geo_cal(ptx[1][1], pty[1][1], ptx[1][0], pty[1][0], nx, ny, ph, &a, &fx, &fy);
            GH(C->aE, i, j) = a;
            GH(C->fcex, i, j) = fx;
            GH(C->fcey, i, j) = fy;
            orc_edge_geo_cal(ptx[1][0], pty[1][0], ptx[0][0], pty[0][0], nx, ny, ph, &a, &fx, &fy);
            GH(C->aS, i, j) = a;
            FW(fcsx, i, j) = fx;
            FW(fcsy, i, j) = fy;
        }
    /* wall length from the solid edge fractions (:79-89) */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            const double ex = (GH(C->aE, i, j) - GH(C->aW, i, j)) * G->dx[i - 1];
            const double ny = (GH(C->aN, i, j) - GH(C->aS, i, j)) * G->dy[j - 1];
            CL(C->WlLh, i, j) = sqrt(ex * ex + ny * ny);
        }
    /* face-centre ghosts from the raw edge centres (:91-117); the corner
     * entries of the first four statements read rows 0 and nj+1 before any
     * statement defines them and are overwritten by the last four */
    const double dx1 = G->dx[0], dxn = G->dx[ni - 1], dy1 = G->dy[0], dyn = G->dy[nj - 1];
    for (int j = 1; j <= nj; ++j) {
        GH(C->fcex, 0, j) = GH(C->fcex, 1, j) - dx1;
        GH(C->fcey, 0, j) = GH(C->fcey, 1, j);
        GH(C->fcex, ni + 1, j) = GH(C->fcex, ni, j) + dxn;
        GH(C->fcey, ni + 1, j) = GH(C->fcey, ni, j);
        GH(C->fcnx, 0, j) = GH(C->fcnx, 1, j) - dx1;
        GH(C->fcny, 0, j) = GH(C->fcny, 1, j);
        GH(C->fcnx, ni + 1, j) = GH(C->fcnx, ni, j) + dxn;
        GH(C->fcny, ni + 1, j) = GH(C->fcny, ni, j);
    }
    for (int i = 0; i <= ni + 1; ++i) {
        GH(C->fcex, i, 0) = GH(C->fcex, i, 1);
        GH(C->fcey, i, 0) = GH(C->fcey, i, 1) - dy1;
        GH(C->fcex, i, nj + 1) = GH(C->fcex, i, nj);
        GH(C->fcey, i, nj + 1) = GH(C->fcey, i, nj) + dyn;
    }
    /* the FCN ghost rows copy FCE (:112-117) */
    for (int i = 0; i <= ni + 1; ++i) {
        GH(C->fcnx, i, 0) = GH(C->fcex, i, 1);
        GH(C->fcny, i, 0) = GH(C->fcey, i, 1) - dy1;
        GH(C->fcnx, i, nj + 1) = GH(C->fcex, i, nj);
        GH(C->fcny, i, nj + 1) = GH(C->fcey, i, nj) + dyn;
    }
    /* solid -> fluid edge fractions and fluid face centres (:119-159) */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            const double dx = G->dx[i - 1], dy = G->dy[j - 1];
            GH(C->aW, i, j) = 1.0 - GH(C->aW, i, j);
            GH(C->aE, i, j) = 1.0 - GH(C->aE, i, j);
            GH(C->aN, i, j) = 1.0 - GH(C->aN, i, j);
            GH(C->aS, i, j) = 1.0 - GH(C->aS, i, j);
            if (fabs(GH(C->fcey, i, j)) < EPSI)
                GH(C->fcey, i, j) = 0.0;
            else
                GH(C->fcey, i, j) = (0.5 - GH(C->aE, i, j) / 2.0) * copysign(1.0, -GH(C->fcey, i, j)) * dy;
            if (GH(C->aE, i, j) < EPSIF) GH(C->fcey, i, j) = 0.0;
            if (fabs(FW(fcwy, i, j)) < EPSI)
                FW(fcwy, i, j) = 0.0;
            else
                FW(fcwy, i, j) = (0.5 - GH(C->aW, i, j) / 2.0) * copysign(1.0, -FW(fcwy, i, j)) * dy;
            if (GH(C->aW, i, j) < EPSIF) FW(fcwy, i, j) = 0.0;
            if (fabs(GH(C->fcnx, i, j)) < EPSI)
                GH(C->fcnx, i, j) = 0.0;
            else
                GH(C->fcnx, i, j) = (0.5 - GH(C->aN, i, j) / 2.0) * copysign(1.0, -GH(C->fcnx, i, j)) * dx;
            if (GH(C->aN, i, j) < EPSIF) GH(C->fcnx, i, j) = 0.0;
            if (fabs(FW(fcsx, i, j)) < EPSI)
                FW(fcsx, i, j) = 0.0;
            else
                FW(fcsx, i, j) = (0.5 - GH(C->aS, i, j) / 2.0) * copysign(1.0, -FW(fcsx, i, j)) * dx;
            if (GH(C->aS, i, j) < EPSIF) FW(fcsx, i, j) = 0.0;
        }
    /* harmonic face averaging and area-weighted face centres (:161-204) */
    for (int i = 2; i <= ni - 1; ++i)
        for (int j = 2; j <= nj - 1; ++j) {
            double n_ = GH(C->aN, i, j), s_ = GH(C->aS, i, j + 1);
            double avg = 2.0 * n_ * s_ / (n_ + s_ + ROUNDOFF);
            GH(C->fcnx, i, j) = (GH(C->fcnx, i, j) * n_ + FW(fcsx, i, j + 1) * s_) / (n_ + s_ + ROUNDOFF);
            GH(C->fcny, i, j) = (GH(C->fcny, i, j) * n_ + (FW(fcsy, i, j + 1) + (G->y[j] - G->y[j - 1])) * s_) /
                                (n_ + s_ + ROUNDOFF);
            if (n_ < EPSIF && s_ < EPSIF) GH(C->fcny, i, j) = G->dy[j - 1] / 2.0;
            GH(C->aN, i, j) = avg;
            GH(C->aS, i, j + 1) = avg;
            const double e_ = GH(C->aE, i, j), w_ = GH(C->aW, i + 1, j);
            avg = 2.0 * e_ * w_ / (e_ + w_ + ROUNDOFF);
            GH(C->fcex, i, j) = (GH(C->fcex, i, j) * e_ + (FW(fcwx, i + 1, j) + (G->x[i] - G->x[i - 1])) * w_) /
                                (e_ + w_ + ROUNDOFF);
            GH(C->fcey, i, j) = (GH(C->fcey, i, j) * e_ + FW(fcwy, i + 1, j) * w_) / (e_ + w_ + ROUNDOFF);
            if (e_ < EPSIF && w_ < EPSIF) GH(C->fcex, i, j) = G->dx[i - 1] / 2.0;
            GH(C->aE, i, j) = avg;
            GH(C->aW, i + 1, j) = avg;
        }
    /* solid cells close their faces and their neighbours'; on P cells a
     * cell shut on both sides of an axis becomes solid (:214-243), visited
     * in the reference's order: each cell sees the closings of the cells
     * visited before it */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            if (CL(C->vofS, i, j) > 1.0 - EPSI) close_cell(ni, C, i, j);
            if ((GH(C->aW, i, j) + GH(C->aE, i, j) < ROUNDOFF || GH(C->aS, i, j) + GH(C->aN, i, j) < ROUNDOFF) &&
                CL(C->vofS, i, j) < 1.0 - EPSI && p_cell)
                close_cell(ni, C, i, j);
        }
    /* wall length (and delh) of full solid cells (:245-258) */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            if (fabs(CL(C->vofS, i, j) - 1.0) < ROUNDOFF) {
                if (GH(C->aW, i, j) == 0.0 || GH(C->aE, i, j) == 0.0) {
                    CL(C->WlLh, i, j) = G->dy[j - 1];
                    if (C->delh) CL(C->delh, i, j) = G->dx[i - 1] / 2.0;
                }
                if (GH(C->aN, i, j) == 0.0 || GH(C->aS, i, j) == 0.0) {
                    CL(C->WlLh, i, j) = G->dx[i - 1];
                    if (C->delh) CL(C->delh, i, j) = G->dy[j - 1] / 2.0;
                }
            }
        }
    /* corner patch under the boom's bar and under flat solid bottoms
     * (ICorProb, :259-286).  The reference writes the north aperture from
     * the default-real literal 0.40 (0.4000000059604645) and the south one
     * of the cell above from 0.4d0; both are 0.4 here, so the face has one
     * aperture and the Poisson rows on either side of it stay symmetric. */
    if (icorprob && p_cell) {
        for (int i = 1; i <= ni; ++i)
            for (int j = 1; j <= nj - 4; ++j) {
                const double s = CL(C->vofS, i, j);
                if (s > -EPSI && s < EPSI) {
                    const double s1 = CL(C->vofS, i, j + 1), s2 = CL(C->vofS, i, j + 2);
                    if (s1 > 0.6 - EPSI && s1 < 0.6 + EPSI) {
                        if (s2 > 0.6 - EPSI && s2 < 0.6 + EPSI) {
                            GH(C->aN, i, j) = 0.4;
                            GH(C->aS, i, j + 1) = 0.4;
                            GH(C->fcnx, i, j) = -0.3 * G->dx[i - 1];
                            GH(C->fcny, i, j) = 0.5 * G->dy[j - 1];
                            CL(C->WlLh, i, j) = 0.6 * G->dx[i - 1];
                            CL(C->nyS, i, j) = -1.0;
                            CL(C->nxS, i, j) = 0.0;
                        }
                    }
                    if (s1 > 1.0 - EPSI && s1 < 1.0 + EPSI) {
                        CL(C->WlLh, i, j) = G->dx[i - 1];
                        CL(C->nyS, i, j) = -1.0;
                        CL(C->nxS, i, j) = 0.0;
                    }
                }
            }
    }
#undef FW
    free(fcwx);
    free(fcwy);
    free(fcsx);
    free(fcsy);
}

/* Cutcell_Mod.f90:360-397 */
void orc_define_moex(int ni, int nj, orc_ccell* P, orc_ccell* U, orc_ccell* V) {
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            CL(U->moex, i, j) = 0;
            CL(V->moex, i, j) = 0;
            const double su = CL(U->vofS, i, j), sv = CL(V->vofS, i, j);
            if (su >= EPSI && su < 1.0 - EPSI) {
                if (i < ni) {
                    if (CL(P->vofS, i, j) >= 1.0 - EPSI || CL(P->vofS, i + 1, j) >= 1.0 - EPSI ||
                        GH(P->aE, i, j) < EPSIF) {
                        CL(U->moex, i, j) = 1;
                        GH(P->aE, i, j) = 0.0;
                        GH(P->aW, i + 1, j) = 0.0;
                    }
                } else if (CL(P->vofS, i, j) >= 1.0 - EPSI) {
                    CL(U->moex, i, j) = 1;
                }
            }
            if (sv >= EPSI && sv < 1.0 - EPSI) {
                if (j < nj) {
                    if (CL(P->vofS, i, j) >= 1.0 - EPSI || CL(P->vofS, i, j + 1) >= 1.0 - EPSI ||
                        GH(P->aN, i, j) < EPSIF) {
                        CL(V->moex, i, j) = 1;
                        GH(P->aN, i, j) = 0.0;
                        GH(P->aS, i, j + 1) = 0.0;
                    }
                } else if (CL(P->vofS, i, j) >= 1.0 - EPSI) {
                    CL(V->moex, i, j) = 1;
                }
            }
        }
}

/* Cutcell_Mod.f90:340-358: the running number of each fluid cell over
 * i = 1..Isize-iu (outer), j = 1..Jsize-iv, -1 for the solid ones there;
 * the cells the loop does not reach keep 0 (see the file header). */
void orc_number_external(int ni, int nj, orc_ccell* C, int iu, int iv) {
    int ctr = 0;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) CL(C->posnu, i, j) = 0;
    for (int i = 1; i <= ni - iu; ++i)
        for (int j = 1; j <= nj - iv; ++j) {
            if (CL(C->vofS, i, j) < 1.0 - EPSI) {
                CL(C->posnu, i, j) = ctr;
                ctr = ctr + 1;
            } else {
                CL(C->posnu, i, j) = -1;
            }
        }
}

/* ghosts neither routine writes: the nearest written entry, rows first */
static void fill_unwritten(int ni, int nj, double* a, int east) {
    if (east) {
        for (int j = 1; j <= nj; ++j) GH(a, ni + 1, j) = GH(a, ni, j);
        for (int i = 0; i <= ni + 1; ++i) {
            GH(a, i, 0) = GH(a, i, 1);
            GH(a, i, nj + 1) = GH(a, i, nj);
        }
    } else {
        for (int i = 1; i <= ni; ++i) GH(a, i, nj + 1) = GH(a, i, nj);
        for (int j = 0; j <= nj + 1; ++j) {
            GH(a, 0, j) = GH(a, 1, j);
            GH(a, ni + 1, j) = GH(a, ni, j);
        }
    }
}

/* Cutcell_Mod.f90:551-616; bdx = BGrid%dx(i+iu, .) per i (index i-1 for
 * the reference's i+iu) */
static void east_interpolation(int ni, int nj, orc_ccell* T, const orc_cgrid* TG, const double* bdx) {
    for (int j = 1; j <= nj; ++j) {
        for (int i = 1; i <= ni - 1; ++i) {
            const double bd = bdx[i - 1];
            if (CL(T->posnu, i, j) != -1 && CL(T->posnu, i + 1, j) != -1) {
                GH(T->SxE, i, j) = CL(T->ccx, i + 1, j) + bd - CL(T->ccx, i, j);
                const double Sy = CL(T->ccy, i + 1, j) - CL(T->ccy, i, j);
                GH(T->EtaE, i, j) = fabs(GH(T->fcex, i, j) - CL(T->ccx, i, j)) / GH(T->SxE, i, j);
                if (GH(T->EtaE, i, j) >= 1.0) GH(T->EtaE, i, j) = 0.5;
                const double eta = GH(T->EtaE, i, j);
                const double xf = (1.0 - eta) * CL(T->ccx, i, j) + eta * (CL(T->ccx, i + 1, j) + bd);
                const double yf = (1.0 - eta) * CL(T->ccy, i, j) + eta * CL(T->ccy, i + 1, j);
                if (fabs(Sy) > 0.0001 * TG->dy[j - 1]) {
                    const double nyf = 0.5 * (CL(T->nyS, i, j) + CL(T->nyS, i + 1, j));
                    const double ipd =
                        0.5 * (fabs(xf * CL(T->nxS, i, j) + yf * CL(T->nyS, i, j) + CL(T->phiS, i, j)) +
                               fabs((xf - bd) * CL(T->nxS, i + 1, j) + yf * CL(T->nyS, i + 1, j) +
                                    CL(T->phiS, i + 1, j)));
                    const double fcd =
                        0.5 * (fabs(GH(T->fcex, i, j) * CL(T->nxS, i, j) + GH(T->fcey, i, j) * CL(T->nyS, i, j) +
                                    CL(T->phiS, i, j)) +
                               fabs((GH(T->fcex, i, j) - bd) * CL(T->nxS, i + 1, j) +
                                    GH(T->fcey, i, j) * CL(T->nyS, i + 1, j) + CL(T->phiS, i + 1, j)));
                    GH(T->AlE, i, j) = fabs(fcd) / fabs(ipd);
                    GH(T->DAlE, i, j) = Sy * nyf / (GH(T->SxE, i, j) * fabs(ipd));
                    if (fabs(ipd) < 1.0e-5) {
                        GH(T->DAlE, i, j) = 0.0;
                        GH(T->AlE, i, j) = 1.0;
                    }
                    if (GH(T->AlE, i, j) >= 2.0) GH(T->AlE, i, j) = 2.0;
                } else {
                    GH(T->DAlE, i, j) = 0.0;
                    GH(T->AlE, i, j) = 1.0;
                }
            } else {
                GH(T->AlE, i, j) = 1.0;
                GH(T->DAlE, i, j) = 0.0;
                GH(T->SxE, i, j) = bd;
                GH(T->EtaE, i, j) = 0.5;
            }
        }
        GH(T->SxE, ni, j) = GH(T->SxE, ni - 1, j);
        GH(T->EtaE, ni, j) = GH(T->EtaE, ni - 1, j);
        GH(T->DAlE, ni, j) = GH(T->DAlE, ni - 1, j);
        GH(T->AlE, ni, j) = GH(T->AlE, ni - 1, j);
        GH(T->SxE, 0, j) = GH(T->SxE, 1, j);
        GH(T->EtaE, 0, j) = GH(T->EtaE, 1, j);
        GH(T->DAlE, 0, j) = GH(T->DAlE, 1, j);
        GH(T->AlE, 0, j) = GH(T->AlN, 1, j);
    }
    fill_unwritten(ni, nj, T->SxE, 1);
    fill_unwritten(ni, nj, T->EtaE, 1);
    fill_unwritten(ni, nj, T->DAlE, 1);
    fill_unwritten(ni, nj, T->AlE, 1);
}

/* Cutcell_Mod.f90:618-700; bdy = BGrid%dy(., j+iv) per j */
static void north_interpolation(int ni, int nj, orc_ccell* T, const orc_cgrid* TG, const double* bdy) {
    for (int i = 1; i <= ni; ++i) {
        for (int j = 1; j <= nj - 1; ++j) {
            const double bd = bdy[j - 1];
            if (CL(T->posnu, i, j) != -1 && CL(T->posnu, i, j + 1) != -1) {
                GH(T->SyN, i, j) = CL(T->ccy, i, j + 1) + bd - CL(T->ccy, i, j);
                const double Sx = CL(T->ccx, i, j + 1) - CL(T->ccx, i, j);
                GH(T->EtaN, i, j) = fabs(GH(T->fcny, i, j) - CL(T->ccy, i, j)) / GH(T->SyN, i, j);
                if (GH(T->EtaN, i, j) >= 1.0) GH(T->EtaN, i, j) = 0.5;
                const double eta = GH(T->EtaN, i, j);
                const double xf = (1.0 - eta) * CL(T->ccx, i, j) + eta * CL(T->ccx, i, j + 1);
                const double yf = (1.0 - eta) * CL(T->ccy, i, j) + eta * (CL(T->ccy, i, j + 1) + bd);
                if (fabs(Sx) >= 0.0001 * TG->dx[i - 1]) {
                    const double nxf = 0.5 * (CL(T->nxS, i, j) + CL(T->nxS, i, j + 1));
                    const double ipd =
                        0.5 * (fabs(xf * CL(T->nxS, i, j) + yf * CL(T->nyS, i, j) + CL(T->phiS, i, j)) +
                               fabs(xf * CL(T->nxS, i, j + 1) + (yf - bd) * CL(T->nyS, i, j + 1) +
                                    CL(T->phiS, i, j + 1)));
                    const double fcd =
                        0.5 * (fabs(GH(T->fcnx, i, j) * CL(T->nxS, i, j) + GH(T->fcny, i, j) * CL(T->nyS, i, j) +
                                    CL(T->phiS, i, j)) +
                               fabs(GH(T->fcnx, i, j) * CL(T->nxS, i, j + 1) +
                                    (GH(T->fcny, i, j) - bd) * CL(T->nyS, i, j + 1) + CL(T->phiS, i, j + 1)));
                    GH(T->AlN, i, j) = fabs(ipd) / fabs(fcd);
                    GH(T->DAlN, i, j) = Sx * nxf / (GH(T->SyN, i, j) * fabs(ipd));
                    if (fabs(ipd) < 1.0e-5) {
                        GH(T->DAlN, i, j) = 0.0;
                        GH(T->AlN, i, j) = 1.0;
                    }
                    if (GH(T->AlN, i, j) >= 2.0) GH(T->AlN, i, j) = 2.0;
                } else {
                    GH(T->DAlN, i, j) = 0.0;
                    GH(T->AlN, i, j) = 1.0;
                }
            } else {
                GH(T->EtaN, i, j) = 0.5;
                GH(T->AlN, i, j) = 1.0;
                GH(T->DAlN, i, j) = 0.0;
                GH(T->SyN, i, j) = bd;
            }
        }
        GH(T->SyN, i, nj) = GH(T->SyN, i, nj - 1);
        GH(T->EtaN, i, nj) = GH(T->EtaN, i, nj - 1);
        GH(T->DAlN, i, nj) = GH(T->DAlN, i, nj - 1);
        GH(T->AlN, i, nj) = GH(T->AlN, i, nj - 1);
        GH(T->SyN, i, 0) = GH(T->SyN, i, 1);
        GH(T->EtaN, i, 0) = GH(T->EtaN, i, 1);
        GH(T->DAlN, i, 0) = GH(T->DAlN, i, 1);
        GH(T->AlN, i, 0) = GH(T->AlN, i, 1);
    }
    fill_unwritten(ni, nj, T->SyN, 0);
    fill_unwritten(ni, nj, T->EtaN, 0);
    fill_unwritten(ni, nj, T->DAlN, 0);
    fill_unwritten(ni, nj, T->AlN, 0);
}

/* Cutcell_Mod.f90:399-549 + the four interpolation calls (:545-548).  The
 * staggered cells' centroid offsets (Cell_Cent), delh, WlLh and face
 * factors; reads the P cells' face centres FCE / FCN. */
void orc_new_cell_face(const orc_cgrid* PG, const orc_cgrid* UG, const orc_cgrid* VG, orc_ccell* P, orc_ccell* U,
                       orc_ccell* V) {
    const int ni = PG->ni, nj = PG->nj;
    const double tol = 1.0e-24;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            if (CL(U->moex, i, j) != 1 && CL(U->vofS, i, j) > EPSI) {
                CL(U->ccx, i, j) = GH(P->fcex, i, j) - PG->dx[i - 1] / 2.0;
                CL(U->ccy, i, j) = GH(P->fcey, i, j);
            }
            if (CL(V->moex, i, j) != 1 && CL(V->vofS, i, j) > EPSI) {
                CL(V->ccx, i, j) = GH(P->fcnx, i, j);
                CL(V->ccy, i, j) = GH(P->fcny, i, j) - PG->dy[j - 1] / 2.0;
            }
        }
    /* very small cells connected to one P cell (:427-489) */
    orc_ccell* T2[2] = {U, V};
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            for (int k = 0; k < 2; ++k) {
                orc_ccell* T = T2[k];
                if (CL(T->moex, i, j) != 1 || !(CL(T->vofS, i, j) < 1.0 - EPSI)) continue;
                const double nN = GH(T->aN, i, j), nS = GH(T->aN, i, j - 1);
                const double eE = GH(T->aE, i, j), eW = GH(T->aE, i - 1, j);
                if (nN * nS == 0.0) {
                    CL(T->ccx, i, j) = GH(T->fcnx, i, j) * nN / (nN + tol) + GH(T->fcnx, i, j - 1) * nS / (nS + tol);
                } else {
                    CL(T->ccy, i, j) = GH(T->fcey, i, j) * eE / (eE + tol) + GH(T->fcey, i - 1, j) * eW / (eW + tol);
                    if (k == 0)
                        CL(T->ccx, i, j) = -(CL(T->nyS, i, j) * CL(T->ccy, i, j) + CL(T->phiS, i, j)) /
                                           (CL(T->nxS, i, j) + tol);
                    else
                        CL(T->ccx, i, j) = -(CL(T->ccy, i, j) * CL(T->nyS, i, j) + CL(T->phiS, i, j)) /
                                           (CL(T->nxS, i, j) + tol);
                }
                if (eE * eW == 0.0) {
                    CL(T->ccy, i, j) = GH(T->fcey, i, j) * eE / (eE + tol) + GH(T->fcey, i - 1, j) * eW / (eW + tol);
                } else {
                    if (k == 0)
                        CL(T->ccy, i, j) = -(CL(T->nxS, i, j) * CL(T->ccx, i, j) + CL(T->phiS, i, j)) /
                                           (CL(T->nyS, i, j) + tol);
                    else
                        CL(T->ccy, i, j) = -(CL(T->ccx, i, j) * CL(T->nxS, i, j) + CL(T->phiS, i, j)) /
                                           (CL(T->nyS, i, j) + tol);
                }
            }
    /* distance to the wall and wall length from the final apertures
     * (:490-543) */
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j)
            for (int k = 0; k < 2; ++k) {
                orc_ccell* T = T2[k];
                const orc_cgrid* TG = k == 0 ? UG : VG;
                CL(T->delh, i, j) = fabs(CL(T->ccx, i, j) * CL(T->nxS, i, j) + CL(T->ccy, i, j) * CL(T->nyS, i, j) +
                                         CL(T->phiS, i, j)) +
                                    tol;
                if (CL(T->moex, i, j) == 1) CL(T->delh, i, j) = 1.0e-20;
                const double ex = (GH(T->aE, i, j) - GH(T->aW, i, j)) * TG->dx[i - 1];
                const double ny = (GH(T->aN, i, j) - GH(T->aS, i, j)) * TG->dy[j - 1];
                CL(T->WlLh, i, j) = sqrt(ex * ex + ny * ny);
            }
    /* BGrid spacings: U east -> PGrid%dx(i+1), U north -> VGrid%dy(j),
     * V east -> UGrid%dx(i), V north -> PGrid%dy(j+1) */
    double* bx = (double*)malloc((size_t)ni * sizeof(double));
    double* by = (double*)malloc((size_t)nj * sizeof(double));
    for (int i = 1; i <= ni; ++i) bx[i - 1] = i < ni ? PG->dx[i] : PG->dx[ni - 1];
    east_interpolation(ni, nj, U, UG, bx);
    north_interpolation(ni, nj, U, UG, VG->dy);
    east_interpolation(ni, nj, V, VG, UG->dx);
    for (int j = 1; j <= nj; ++j) by[j - 1] = j < nj ? PG->dy[j] : PG->dy[nj - 1];
    north_interpolation(ni, nj, V, VG, by);
    free(bx);
    free(by);
}

/* Grid_Preprocess (Cutcell_Mod.f90:18-32) then NewCellFace, as
 * Solver_Mod.f90:303-304 call them.  P, U, V hold the solid fields on
 * entry (vofS from frac of the solid line, Clsvof_Mod.f90:273-277) and the
 * AlN of the previous call (read by AlE(0,j)). */
void orc_cut_cell_geometry(const orc_cgrid* PG, const orc_cgrid* UG, const orc_cgrid* VG, orc_ccell* P,
                           orc_ccell* U, orc_ccell* V, int icorprob) {
    const int ni = PG->ni, nj = PG->nj;
    orc_cell_geo_cal(PG, P, icorprob);
    orc_cell_geo_cal(UG, U, icorprob);
    orc_cell_geo_cal(VG, V, icorprob);
    orc_define_moex(ni, nj, P, U, V);
    orc_number_external(ni, nj, P, 0, 0);
    orc_number_external(ni, nj, U, 1, 0);
    orc_number_external(ni, nj, V, 0, 1);
    orc_new_cell_face(PG, UG, VG, P, U, V);
}
