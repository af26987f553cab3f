/*
 * CPU FP64 oracle of the time-step control around the hot path.
 * TEST INFRASTRUCTURE ONLY (see sg_oracle.h).
 *
 * orc_time_step restates ComputeTimeStep (Solver_Mod.f90:321-344);
 * orc_closed_box_bc is the closed-box form of Boundary_Condition_Var
 * (StateVariables.f90:118-211) used by the box cases of this build (the
 * reference's routine imposes its wave channel's inlet/outlet profiles; the
 * closed box keeps its copy rules: pressure and tangential velocity copied
 * into the ghost layer, p = 0 above the open top, walls u(0,j) = u(ni,j) = 0
 * and v(i,0) = 0).
 */
#include <math.h>
#include <stddef.h>

#include "sg_oracle.h"

#define G(a, i, j) (a)[(size_t)(i) + (size_t)(j) * (size_t)(ni + 2)]

/* Solver_Mod.f90:321-344; u, v ghosted TVar%u/v; udx = UGrid%dx(:,1),
 * vdy = VGrid%dy(1,:); vint, uref = TVar%Vint, TVar%Uref; vs, asy the
 * solid's velocity and acceleration (0 for solids at rest). */
double orc_time_step(int ni, int nj, const double* udx, const double* vdy, const double* u, const double* v,
                     double cfl, double Fr, double vint, double uref, double vs, double asy) {
    double dt = 1.0;
    for (int j = 1; j <= nj; ++j) {
        for (int i = 1; i <= ni; ++i) {
            const double dx = udx[i - 1], dy = vdy[j - 1];
            const double uu = G(u, i, j), vv = G(v, i, j);
            dt = fmin(dt, cfl * dx / fabs(uu));
            dt = fmin(dt, cfl * dy / fabs(vv));
            dt = fmin(fmin(dt, 2.0 * cfl * dy / (fabs(vv) + sqrt(vv * vv + 4.0 * dy / Fr))), cfl * dy / (vint / uref));
            dt = fmin(dt, cfl * dy / (vint / uref));
            dt = fmin(dt, cfl * dy / fabs(vs));
            dt = fmin(dt, (-vs + sqrt(fmax(fabs(vs * vs + 2.0 * fabs(asy) * 0.05 * dy), 1e-30))) / fabs(asy + 1.e-20));
        }
        dt = fmin(dt, cfl * udx[0] / fabs(G(u, 0, j)));
    }
    return dt;
}

void orc_closed_box_bc(int ni, int nj, double* u, double* v, double* p) {
    for (int j = 1; j <= nj; ++j) {
        G(p, 0, j) = G(p, 1, j);
        G(p, ni + 1, j) = G(p, ni, j);
        G(u, 0, j) = 0.0;
        G(u, ni, j) = 0.0;
        G(u, ni + 1, j) = 0.0;
        G(v, 0, j) = G(v, 1, j);
        G(v, ni + 1, j) = G(v, ni, j);
    }
    for (int i = 1; i <= ni; ++i) {
        G(p, i, 0) = G(p, i, 1);
        G(u, i, 0) = G(u, i, 1);
        G(v, i, 0) = 0.0;
        G(p, i, nj + 1) = 0.0;
        G(u, i, nj + 1) = G(u, i, nj);
        G(v, i, nj + 1) = G(v, i, nj);
    }
}
