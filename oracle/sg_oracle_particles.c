/*
 * CPU FP64 oracle of the Lagrangian particle tracker.
 * TEST INFRASTRUCTURE ONLY (see sg_oracle.h).
 *
 * orc_particle_track restates TrackingParticles (Particles.f90:71-205) with
 * Drag (:206-223) and ParticlePosition (:228-249), in the reference's
 * expression order.  The fluid constants (row, roa, nuw, nua;
 * StateVariables.f90:26-27) are parameters, as everywhere in this build.
 * Quirks kept: the fluid velocity at a particle is interpolated at its
 * position at the start of the time step for all itp = 10 sub-steps (the
 * sub-steps advance xyp, not Posp); u is interpolated linearly in x only, v
 * in y only; a particle outside the grid is left alone; a particle past the
 * last cell centre, or in a cell that is neither fluid nor full solid, is
 * parked at (1e3, 1e3) with its velocity unchanged.  The reference's
 * `pause` on a NaN or runaway particle becomes a count returned to the
 * caller.
 */
#include <math.h>
#include <stddef.h>

#include "sg_oracle.h"

#define ITP 10
#define TOLP 1.0e-20
#define CA 0.5
#define EPSIP 1.0e-3

/* Particles.f90:206-223 (Schiller-Naumann) */
double orc_particle_drag(double re) {
    if (re < 1.0) return 24.0 / (re + TOLP);
    if (re <= 1000.0) return 24.0 / (re + TOLP) * (1.0 + 0.15 * pow(re, 0.687));
    return 0.44;
}

/* Particles.f90:228-249: first column / row whose closed interval holds the
 * point (1-based), -1 if none */
void orc_particle_position(int ni, int nj, const double* xc, const double* yc, const double* dx, const double* dy,
                           double lref, double x, double y, int* ii, int* jj) {
    *ii = -1;
    for (int i = 1; i <= ni; ++i) {
        if (x / lref >= xc[i - 1] - dx[i - 1] / 2.0 && x / lref <= xc[i - 1] + dx[i - 1] / 2.0) {
            *ii = i;
            break;
        }
    }
    *jj = -1;
    for (int j = 1; j <= nj; ++j) {
        if (y / lref >= yc[j - 1] - dy[j - 1] / 2.0 && y / lref <= yc[j - 1] + dy[j - 1] / 2.0) {
            *jj = j;
            break;
        }
    }
}

#define GU(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(ni + 2)]
#define CL(a, i, j) (a)[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)ni]

/* Particles.f90:71-205 for n particles: px, py, pu, pv in/out, pd the
 * diameters; vof, vofS cell arrays; u, v ghosted TVar%u/v; dt the
 * non-dimensional time step.  Returns the number of particles that met the
 * reference's failure test (NaN velocity, |x| > 1e10 or |ax| > 1e10). */
int orc_particle_track(int ni, int nj, const double* xc, const double* yc, const double* dx, const double* dy,
                       const double* vof, const double* vofS, const double* u, const double* v, int n, double* px,
                       double* py, double* pu, double* pv, const double* pd, const orc_particle_params* P,
                       double dt) {
    int bad = 0;
    const double dtp = dt * (P->lref / P->uref) / (double)ITP;
    for (int p = 0; p < n; ++p) {
        double xx = px[p], yy = py[p], upo = pu[p], vpo = pv[p];
        const double dpp = pd[p];
        int flagged = 0;
        for (int itep = 1; itep <= ITP; ++itep) {
            int ii, jj;
            orc_particle_position(ni, nj, xc, yc, dx, dy, P->lref, px[p], py[p], &ii, &jj);
            if (ii == -1 || jj == -1) continue;
            const double f = CL(vof, ii, jj), s = CL(vofS, ii, jj);
            if (px[p] / P->lref < xc[ni - 1] && py[p] / P->lref < yc[nj - 1] && s < 1.0 - EPSIP) {
                const double ropp = P->rho_l * f / (1.0 - s + TOLP) + P->rho_g * (1.0 - f - s) / (1.0 - s + TOLP);
                const double gama = P->rho_p / ropp;
                const double nupp = P->nu_l * f / (1.0 - s + TOLP) + P->nu_g * (1.0 - f - s) / (1.0 - s + TOLP);
                const double dudx = (GU(u, ii, jj) - GU(u, ii - 1, jj)) / dx[ii - 1];
                const double dvdy = (GU(v, ii, jj) - GU(v, ii, jj - 1)) / dy[jj - 1];
                const double ug0 = 0.5 * (GU(u, ii, jj) + GU(u, ii - 1, jj));
                const double vg0 = 0.5 * (GU(v, ii, jj) + GU(v, ii, jj - 1));
                const double ug = (ug0 + dudx * (px[p] / P->lref - xc[ii - 1])) * P->uref;
                const double vg = (vg0 + dvdy * (py[p] / P->lref - yc[jj - 1])) * P->uref;
                double vrel = sqrt((upo - ug) * (upo - ug) + (vpo - vg) * (vpo - vg));
                if (vrel < 1.0e-7) vrel = 1.0e-7;
                const double reyp = dpp * vrel / nupp;
                const double cd = orc_particle_drag(reyp);
                const double mag = sqrt((upo - ug) * (upo - ug) + (vpo - vg) * (vpo - vg));
                const double axp = (1.0 - gama) / (gama + CA) * P->gx -
                                   3.0 * cd / 8.0 / (dpp / 2.0) / (gama + CA) * (upo - ug) * mag;
                double sig = (1.0 - gama) / (gama + CA) * P->gx;
                const double beta = 3.0 * cd / 8.0 / (dpp / 2.0) / (gama + CA) * mag;
                const double e = exp(-beta * dtp);
                const double upn = upo * e + ug * (1.0 - e) + sig / beta * (1.0 - e);
                xx = xx + 0.5 * (upn + upo) * dtp;
                sig = (1.0 - gama) / (gama + CA) * P->gy;
                const double vpn = vpo * e + vg * (1.0 - e) + sig / beta * (1.0 - e);
                yy = yy + 0.5 * (vpn + vpo) * dtp;
                if (isnan(upo) || fabs(xx) > 1.0e10 || fabs(axp) > 1.0e10 || isnan(upn) || isnan(vpn)) flagged = 1;
                upo = upn;
                vpo = vpn;
            } else if (s > 1.0 - EPSIP) {
                const double upn = P->us, vpn = P->vs;
                xx = xx + 0.5 * (upn + upo) * dtp;
                yy = yy + 0.5 * (vpn + vpo) * dtp;
                upo = upn;
                vpo = vpn;
            } else {
                xx = 1.0e3;
                yy = 1.0e3;
            }
        }
        px[p] = xx;
        py[p] = yy;
        pu[p] = upo;
        pv[p] = vpo;
        bad += flagged;
    }
    return bad;
}
