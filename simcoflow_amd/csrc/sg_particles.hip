// Lagrangian particles on gfx950: TrackingParticles (Particles.f90:71-205)
// with Drag (:206-223) and ParticlePosition (:228-249), one thread per
// particle, FP64, in the reference's expression order with FP contraction
// off (exp and pow are the device library's, so results agree with a CPU
// restatement to rounding, not bit for bit).
//
// Kept from the reference: the fluid velocity at a particle is taken in its
// P-cell at the position of the start of the time step for all itp = 10
// sub-steps (the sub-steps advance the new position, not the one used for the
// lookup); u is interpolated linearly in x only and v in y only; a particle
// outside the grid is left alone; one past the last cell centre, or in a cell
// that is neither fluid nor full solid, is parked at (1e3, 1e3) with its
// velocity unchanged; one in a full-solid cell moves with the solid.
// ParticlePosition scans the columns and rows for the first closed interval
// [x_i - dx_i/2, x_i + dx_i/2] holding the point; here a binary search finds
// the first column whose upper edge is not below the point and the scan goes
// on from there, which picks the same column.
#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int ITP = 10;          // itp, Particles.f90:8
constexpr double TOLP = 1.0e-20; // tol, Particles.f90:9
constexpr double CAP = 0.5;      // Ca, Particles.f90:9

struct PGeo {
    int ni, nj;
    const double *xc, *yc, *dx, *dy;
    const double *vof, *vofS;  // cell arrays
    const double *u, *v;       // ghosted TVar%u, TVar%v
};

__device__ double drag(double re) {
    if (re < 1.0) return 24.0 / (re + TOLP);
    if (re <= 1000.0) return 24.0 / (re + TOLP) * (1.0 + 0.15 * pow(re, 0.687));
    return 0.44;
}

// 1-based index of the first interval [c_k - h_k/2, c_k + h_k/2] holding t, -1 if none
__device__ int locate(const double* __restrict__ c, const double* __restrict__ h, int n, double t) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (c[mid] + h[mid] / 2.0 >= t)
            hi = mid;
        else
            lo = mid + 1;
    }
    for (int k = lo; k < n; ++k) {
        const double a = c[k] - h[k] / 2.0;
        if (t >= a && t <= c[k] + h[k] / 2.0) return k + 1;
        if (a > t) break;
    }
    return -1;
}

__global__ void __launch_bounds__(256) k_particles(PGeo G, int n, double* __restrict__ px, double* __restrict__ py,
                                                   double* __restrict__ pu, double* __restrict__ pv,
                                                   const double* __restrict__ pd, sg_particle_params P, double dt,
                                                   unsigned* __restrict__ flagged) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int ni = G.ni, nj = G.nj;
    const double x0 = px[p], y0 = py[p], dpp = pd[p];
    const double dtp = dt * (P.lref / P.uref) / (double)ITP;
    // the lookup uses the start-of-step position in every sub-step
    const int ii = locate(G.xc, G.dx, ni, x0 / P.lref);
    const int jj = locate(G.yc, G.dy, nj, y0 / P.lref);
    double xx = x0, yy = y0, upo = pu[p], vpo = pv[p];
    bool bad = false;
    if (ii != -1 && jj != -1) {
        const size_t c = (size_t)(ii - 1) + (size_t)(jj - 1) * ni;
        const size_t g = (size_t)ii + (size_t)jj * (ni + 2);
        const double f = G.vof[c], s = G.vofS[c];
        const bool fluid = x0 / P.lref < G.xc[ni - 1] && y0 / P.lref < G.yc[nj - 1] && s < 1.0 - EPSI;
        // the fluid velocity at the particle is the same in every sub-step
        double ug = 0.0, vg = 0.0, gama = 0.0, nupp = 0.0;
        if (fluid) {
            const double ropp = P.rho_l * f / (1.0 - s + TOLP) + P.rho_g * (1.0 - f - s) / (1.0 - s + TOLP);
            gama = P.rho_p / ropp;
            nupp = P.nu_l * f / (1.0 - s + TOLP) + P.nu_g * (1.0 - f - s) / (1.0 - s + TOLP);
            const double dudx = (G.u[g] - G.u[g - 1]) / G.dx[ii - 1];
            const double dvdy = (G.v[g] - G.v[g - (ni + 2)]) / G.dy[jj - 1];
            const double ug0 = 0.5 * (G.u[g] + G.u[g - 1]);
            const double vg0 = 0.5 * (G.v[g] + G.v[g - (ni + 2)]);
            ug = (ug0 + dudx * (x0 / P.lref - G.xc[ii - 1])) * P.uref;
            vg = (vg0 + dvdy * (y0 / P.lref - G.yc[jj - 1])) * P.uref;
        }
        for (int itep = 1; itep <= ITP; ++itep) {
            if (fluid) {
                double vrel = sqrt((upo - ug) * (upo - ug) + (vpo - vg) * (vpo - vg));
                if (vrel < 1.0e-7) vrel = 1.0e-7;
                const double reyp = dpp * vrel / nupp;
                const double cd = drag(reyp);
                const double mag = sqrt((upo - ug) * (upo - ug) + (vpo - vg) * (vpo - vg));
                const double axp = (1.0 - gama) / (gama + CAP) * P.gx -
                                   3.0 * cd / 8.0 / (dpp / 2.0) / (gama + CAP) * (upo - ug) * mag;
                double sig = (1.0 - gama) / (gama + CAP) * P.gx;
                const double beta = 3.0 * cd / 8.0 / (dpp / 2.0) / (gama + CAP) * mag;
                const double e = exp(-beta * dtp);
                const double upn = upo * e + ug * (1.0 - e) + sig / beta * (1.0 - e);
                xx = xx + 0.5 * (upn + upo) * dtp;
                sig = (1.0 - gama) / (gama + CAP) * P.gy;
                const double vpn = vpo * e + vg * (1.0 - e) + sig / beta * (1.0 - e);
                yy = yy + 0.5 * (vpn + vpo) * dtp;
                if (isnan(upo) || fabs(xx) > 1.0e10 || fabs(axp) > 1.0e10 || isnan(upn) || isnan(vpn)) bad = true;
                upo = upn;
                vpo = vpn;
            } else if (s > 1.0 - EPSI) {
                const double upn = P.us, vpn = P.vs;
                xx = xx + 0.5 * (upn + upo) * dtp;
                yy = yy + 0.5 * (vpn + vpo) * dtp;
                upo = upn;
                vpo = vpn;
            } else {
                xx = 1.0e3;
                yy = 1.0e3;
            }
        }
    }
    px[p] = xx;
    py[p] = yy;
    pu[p] = upo;
    pv[p] = vpo;
    if (bad) atomicAdd(flagged, 1u);
}

}  // namespace

int particles_reserve(Ctx* c, int n) {
    ParticleState& S = c->part;
    if (n <= S.cap) return SG_OK;
    // grows only; earlier buffers are released with the context
    const size_t b = (size_t)n * sizeof(double);
    S.x = (double*)dev_alloc(c, b);
    S.y = (double*)dev_alloc(c, b);
    S.u = (double*)dev_alloc(c, b);
    S.v = (double*)dev_alloc(c, b);
    S.d = (double*)dev_alloc(c, b);
    if (!S.flagged) S.flagged = (unsigned*)dev_alloc(c, sizeof(unsigned));
    if (!S.x || !S.y || !S.u || !S.v || !S.d || !S.flagged) {
        set_error("particle buffers: device allocation failed");
        return SG_ERR_ALLOC;
    }
    S.cap = n;
    return SG_OK;
}

int particles_launch(Ctx* c, double dt, const sg_particle_params* prm) {
    ParticleState& S = c->part;
    if (S.n == 0) return SG_OK;
    const PGeo G{c->ni,
                 c->nj,
                 c->cls.xc,
                 c->cls.yc,
                 c->dx,
                 c->dy,
                 (const double*)c->field[SG_F_VOF],
                 (const double*)c->field[SG_F_VOFS],
                 (const double*)c->field[SG_F_U],
                 (const double*)c->field[SG_F_V]};
    hipStream_t s = c->stream;
    SG_HIP(hipMemsetAsync(S.flagged, 0, sizeof(unsigned), s));
    hipLaunchKernelGGL(k_particles, dim3((unsigned)((S.n + 255) / 256)), dim3(256), 0, s, G, S.n, S.x, S.y, S.u,
                       S.v, (const double*)S.d, *prm, dt, S.flagged);
    SG_HIP(hipGetLastError());
    S.last = *prm;
    S.last_dt = dt;
    S.tracked = true;
    return SG_OK;
}

int particles_track(Ctx* c, double dt, const sg_particle_params* prm, int* flagged) {
    ParticleState& S = c->part;
    if (flagged) *flagged = 0;
    if (S.n == 0) return SG_OK;
    const int rc = particles_launch(c, dt, prm);
    if (rc) return rc;
    unsigned h = 0;
    SG_HIP(hipMemcpyAsync(&h, S.flagged, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    if (flagged) *flagged = (int)h;
    return SG_OK;
}

}  // namespace sg
