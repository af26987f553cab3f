/*
 * CPU FP64 oracle of the force on the solid.
 * TEST INFRASTRUCTURE ONLY (see sg_oracle.h).
 *
 * orc_object_force restates ComputeForceObject (Clsvof_Mod.f90:2400-2549)
 * in its loop and expression order, with ICorProb = .TRUE. (Main.f90:205):
 *   clp: over cut P cells p*WlLh*nyS (the neighbour-averaged pressures the
 *        reference computes first are overwritten before use, :2466-2467);
 *        then over i = 1..Isize, j = 1..Jsize-1 the flat faces under full
 *        solid cells (:2471-2512) and the boom-bar corner case (:2516-2530);
 *   clf: over cut V cells that are not momentum-exchange cells,
 *        (v - vs)/Rey/delh times the P cell's wall length (:2533-2543);
 *   force = clf - clp.
 * Arrays: cell arrays (i-1)+(j-1)*ni; p, v ghosted i+j*(ni+2); aE, aW the P
 * apertures (EEdge_Area, WEdge_Area), ghosted.
 */
#include <stddef.h>

#include "sg_oracle.h"

#define EPSI 1.0e-3
#define EPSIF 1.0e-2

#define CL(a, i, j) (a)[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)ni]
#define GH(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(ni + 2)]

double orc_object_force(int ni, int nj, const double* dx, const double* vofS, const double* nyS, const double* WlLh,
                        const double* aE, const double* aW, const double* p, const double* vofSV, const int* moexV,
                        const double* delhV, const double* v, double vs, double Rey, double* clp_out,
                        double* clf_out) {
    double clp = 0.0;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            const double s = CL(vofS, i, j);
            if (s < 1.0 - EPSI && s > EPSI) clp = clp + GH(p, i, j) * CL(WlLh, i, j) * CL(nyS, i, j);
        }
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj - 1; ++j) {
            const double s = CL(vofS, i, j), sn = CL(vofS, i, j + 1);
            if (s < 1.0 - EPSI && GH(aE, i, j) < EPSIF && GH(aW, i, j) < EPSIF) {
                if (sn < 1.0 + EPSI && sn > 1.0 - EPSI) {
                    const double pres = GH(p, i, j - 1), area = dx[i - 1], nyfp = -1.0;
                    clp = clp + pres * area * nyfp - GH(p, i, j) * area * nyfp;
                }
            }
            if (s > -EPSI && s < EPSI) {
                if (sn < 1.0 + EPSI && sn > 1.0 - EPSI) {
                    const double pres = GH(p, i, j), area = dx[i - 1], nyfp = -1.0;
                    clp = clp + pres * area * nyfp;
                }
                if (sn > 0.6 - EPSI && sn < 0.6 + EPSI && j < nj - 1) {
                    const double s2 = CL(vofS, i, j + 2);
                    if (s2 > 0.6 - EPSI && s2 < 0.6 + EPSI) {
                        const double pres = GH(p, i, j), area = 0.6 * dx[i - 1], nyfp = -1.0;
                        clp = clp + pres * area * nyfp;
                    }
                }
            }
        }
    double clf = 0.0;
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            const double sv = CL(vofSV, i, j);
            if (sv > EPSI && sv < 1.0 - EPSI && CL(moexV, i, j) != 1) {
                const double strf = (GH(v, i, j) - vs) / Rey / CL(delhV, i, j);
                clf = clf + strf * CL(WlLh, i, j);
            }
        }
    if (clp_out) *clp_out = clp;
    if (clf_out) *clf_out = clf;
    return clf - clp;
}

/* InterNewVar (Solver_Mod.f90:360-495) in its loop order: U faces whose U
 * cell was full solid at time n and is not now get u = 0, V faces likewise
 * v = vb; a P cell's north / east face whose open fraction shrank by more
 * than 0.6 (new > old / 0.6, either above 0.5) scales v / u by old/new.  The
 * reference's uleft/uright/vbot/vtop and the MsCe lookups are computed but
 * never used, so they are left out.  u, v ghosted (i + j*(ni+2)), in/out;
 * aE/aN (new) and aEo/aNo (time n) ghosted; vofSU/V, vofSUo/Vo cell arrays. */
void orc_inter_new_var(int ni, int nj, const double* vofSUo, const double* vofSU, const double* vofSVo,
                       const double* vofSV, const double* aEo, const double* aE, const double* aNo,
                       const double* aN, double vb, double* u, double* v) {
    for (int i = 1; i <= ni; ++i)
        for (int j = 1; j <= nj; ++j) {
            if (CL(vofSUo, i, j) > 1.0 - EPSI && CL(vofSU, i, j) < 1.0 - EPSI) GH(u, i, j) = 0.0;
            if (CL(vofSVo, i, j) > 1.0 - EPSI && CL(vofSV, i, j) < 1.0 - EPSI) GH(v, i, j) = vb;
            const double nN = GH(aN, i, j), oN = GH(aNo, i, j);
            if (nN > EPSI && oN > EPSI && (nN > 0.5 || oN > 0.5) && oN / nN < 0.6) GH(v, i, j) = GH(v, i, j) * oN / nN;
            const double nE = GH(aE, i, j), oE = GH(aEo, i, j);
            if (nE > EPSI && oE > EPSI && (nE > 0.5 || oE > 0.5) && oE / nE < 0.6) GH(u, i, j) = GH(u, i, j) * oE / nE;
        }
}
