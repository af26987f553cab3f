// Time-step control around the hot path: the CFL step of ComputeTimeStep
// (Solver_Mod.f90:321-344) as a device min-reduction, and the ghost layer of
// a closed box (the box form of Boundary_Condition_Var,
// StateVariables.f90:118-211: pressure and tangential velocity copied into
// the ghosts, p = 0 above the open top, no flow through the walls).
// Compiled without FP contraction: the bounds round exactly as the
// reference's expressions, so the step equals the oracle's bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <mutex>

#include "sg_internal.hpp"

namespace sg {

namespace {

constexpr int NTS = 256;
constexpr int NDP = 7;             // partials per block of the diagnostics

struct StepIn {
    int ni, nj;
    const double *udx, *vdy, *u, *v;
    double cfl, Fr, vint, uref, vs, asy;
};

__device__ inline size_t gh(int ni, int i, int j) { return (size_t)i + (size_t)j * (size_t)(ni + 2); }

// per-block minimum of the step bounds over the cells of a grid-stride range
__global__ void __launch_bounds__(NTS) k_dt(StepIn a, double* part) {
    __shared__ double red[NTS / 64];
    double dt = 1.0;
    const int n = a.ni * a.nj;
    for (int e = blockIdx.x * NTS + threadIdx.x; e < n; e += gridDim.x * NTS) {
        const int i = e % a.ni + 1, j = e / a.ni + 1;
        const double dx = a.udx[i - 1], dy = a.vdy[j - 1];
        const double uu = a.u[gh(a.ni, i, j)], vv = a.v[gh(a.ni, i, j)];
        dt = fmin(dt, a.cfl * dx / fabs(uu));
        dt = fmin(dt, a.cfl * dy / fabs(vv));
        dt = fmin(dt, 2.0 * a.cfl * dy / (fabs(vv) + sqrt(vv * vv + 4.0 * dy / a.Fr)));
        if (i == 1) {
            // the terms that depend on the row alone, once per row: the
            // minimum is exact and fmin drops a NaN in any order, so where
            // they are taken does not change dt
            dt = fmin(dt, a.cfl * dy / (a.vint / a.uref));
            dt = fmin(dt, a.cfl * dy / fabs(a.vs));
            dt = fmin(dt, (-a.vs + sqrt(fmax(fabs(a.vs * a.vs + 2.0 * fabs(a.asy) * 0.05 * dy), 1e-30))) /
                              fabs(a.asy + 1.e-20));
            dt = fmin(dt, a.cfl * a.udx[0] / fabs(a.u[gh(a.ni, 0, j)]));
        }
    }
    for (int o = 32; o >= 1; o >>= 1) dt = fmin(dt, __shfl_xor(dt, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = dt;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < NTS / 64; ++w) m = fmin(m, red[w]);
        part[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(NTS) k_dt_fin(const double* part, int nb, double* out) {
    __shared__ double red[NTS / 64];
    double m = 1.0;
    for (int b = threadIdx.x; b < nb; b += NTS) m = fmin(m, part[b]);
    for (int o = 32; o >= 1; o >>= 1) m = fmin(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < NTS / 64; ++w) r = fmin(r, red[w]);
        out[0] = r;
    }
}

// one thread per boundary position: k < nj the columns, else the rows
__global__ void __launch_bounds__(NTS) k_box_bc(int ni, int nj, double* __restrict__ u, double* __restrict__ v,
                                                double* __restrict__ p) {
    const int k = blockIdx.x * NTS + threadIdx.x;
    if (k < nj) {
        const int j = k + 1;
        p[gh(ni, 0, j)] = p[gh(ni, 1, j)];
        p[gh(ni, ni + 1, j)] = p[gh(ni, ni, j)];
        u[gh(ni, 0, j)] = 0.0;
        u[gh(ni, ni, j)] = 0.0;
        u[gh(ni, ni + 1, j)] = 0.0;
        v[gh(ni, 0, j)] = v[gh(ni, 1, j)];
        v[gh(ni, ni + 1, j)] = v[gh(ni, ni, j)];
    }
}

__global__ void __launch_bounds__(NTS) k_box_bc_rows(int ni, int nj, double* __restrict__ u, double* __restrict__ v,
                                                     double* __restrict__ p) {
    const int k = blockIdx.x * NTS + threadIdx.x;
    if (k < ni) {
        const int i = k + 1;
        p[gh(ni, i, 0)] = p[gh(ni, i, 1)];
        u[gh(ni, i, 0)] = u[gh(ni, i, 1)];
        v[gh(ni, i, 0)] = 0.0;
        p[gh(ni, i, nj + 1)] = 0.0;
        u[gh(ni, i, nj + 1)] = u[gh(ni, i, nj)];
        v[gh(ni, i, nj + 1)] = v[gh(ni, i, nj)];
    }
}

// NaN-propagating maximum (a NaN in a field shows in the diagnostics)
__device__ inline double nmax(double a, double b) { return (b > a || b != b) ? b : a; }

struct DiagIn {
    int ni, nj;
    const double *dx, *dy, *xc, *yc;
    const double *vof, *vofS, *mres;  // cell arrays
    const double *u, *v;              // ghosted
};

// per-block partial sums (liquid volume, gas volume, gas first moments in x
// and y) and maxima (|mres|, |u|, |v|) over a grid-stride range of cells
__global__ void __launch_bounds__(NTS) k_diag(DiagIn a, double* part, int nb) {
    __shared__ double red[NDP][NTS / 64];
    double v[NDP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const int n = a.ni * a.nj;
    for (int e = blockIdx.x * NTS + threadIdx.x; e < n; e += gridDim.x * NTS) {
        const int i = e % a.ni, j = e / a.ni;
        const double area = a.dx[i] * a.dy[j];
        const double f = a.vof[e], gas = 1.0 - f - a.vofS[e];
        v[0] += f * area;
        v[1] += gas * area;
        v[2] += gas * area * a.xc[i];
        v[3] += gas * area * a.yc[j];
        const size_t g = (size_t)(i + 1) + (size_t)(j + 1) * (size_t)(a.ni + 2);
        v[4] = nmax(v[4], fabs(a.mres[e]));
        v[5] = nmax(v[5], fabs(a.u[g]));
        v[6] = nmax(v[6], fabs(a.v[g]));
    }
    for (int o = 32; o >= 1; o >>= 1) {
        for (int k = 0; k < 4; ++k) v[k] += __shfl_xor(v[k], o, 64);
        for (int k = 4; k < NDP; ++k) v[k] = nmax(v[k], __shfl_xor(v[k], o, 64));
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NDP; ++k) red[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x < NDP) {
        const int k = threadIdx.x;
        double r = red[k][0];
        for (int w = 1; w < NTS / 64; ++w) r = k < 4 ? r + red[k][w] : nmax(r, red[k][w]);
        part[(size_t)k * nb + blockIdx.x] = r;
    }
}

// one block: the partials in block order, then one row of the ring
__global__ void __launch_bounds__(NTS) k_diag_fin(const double* part, int nb, const double* dt, DiagRing ring) {
    __shared__ double red[NDP][NTS / 64];
    double v[NDP] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += NTS)
        for (int k = 0; k < NDP; ++k) v[k] = k < 4 ? v[k] + part[(size_t)k * nb + b] : nmax(v[k], part[(size_t)k * nb + b]);
    for (int o = 32; o >= 1; o >>= 1) {
        for (int k = 0; k < 4; ++k) v[k] += __shfl_xor(v[k], o, 64);
        for (int k = 4; k < NDP; ++k) v[k] = nmax(v[k], __shfl_xor(v[k], o, 64));
    }
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NDP; ++k) red[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[NDP];
        for (int k = 0; k < NDP; ++k) {
            r[k] = red[k][0];
            for (int w = 1; w < NTS / 64; ++w) r[k] = k < 4 ? r[k] + red[k][w] : nmax(r[k], red[k][w]);
        }
        const long long cnt = ring.count[0];
        double* row = ring.rows + (size_t)(cnt % SG_DIAG_RING) * SG_DIAG_COLS;
        row[0] = dt[0];
        row[1] = r[0];
        row[2] = r[1];
        row[3] = r[1] > 0.0 ? r[2] / r[1] : 0.0;
        row[4] = r[1] > 0.0 ? r[3] / r[1] : 0.0;
        row[5] = r[4];
        row[6] = r[5];
        row[7] = r[6];
        ring.count[0] = cnt + 1;
    }
}

}  // namespace

int diagnostics_record(Ctx* c) {
    const DiagIn a{c->ni,
                   c->nj,
                   c->dx,
                   c->dy,
                   c->cls.xc,
                   c->cls.yc,
                   (const double*)c->field[SG_F_VOF],
                   (const double*)c->field[SG_F_VOFS],
                   (const double*)c->field[SG_F_MRES],
                   (const double*)c->field[SG_F_U],
                   (const double*)c->field[SG_F_V]};
    int nb = (c->ni * c->nj + NTS - 1) / NTS;
    if (nb > DIAG_BLOCKS) nb = DIAG_BLOCKS;
    hipLaunchKernelGGL(k_diag, dim3(nb), dim3(NTS), 0, c->stream, a, c->diag.partials, nb);
    hipLaunchKernelGGL(k_diag_fin, dim3(1), dim3(NTS), 0, c->stream, (const double*)c->diag.partials, nb,
                       (const double*)c->step_dev, c->diag.ring);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int diagnostics_read(Ctx* c, int max_rows, double* rows, int* n_rows, long long* n_lost) {
    long long cnt = 0;
    SG_HIP(hipMemcpyAsync(&cnt, c->diag.ring.count, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    long long first = c->diag.read;
    long long lost = 0;
    if (cnt - first > SG_DIAG_RING) {
        lost = cnt - first - SG_DIAG_RING;
        first = cnt - SG_DIAG_RING;
    }
    const long long n = std::min<long long>(cnt - first, (long long)max_rows);
    // the rows are contiguous in the ring up to its end, then from its start
    long long done = 0;
    while (done < n) {
        const long long slot = (first + done) % SG_DIAG_RING;
        const long long len = std::min<long long>(n - done, SG_DIAG_RING - slot);
        SG_HIP(hipMemcpyAsync(rows + done * SG_DIAG_COLS, c->diag.ring.rows + slot * SG_DIAG_COLS,
                              (size_t)len * SG_DIAG_COLS * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        done += len;
    }
    SG_HIP(hipStreamSynchronize(c->stream));
    c->diag.read = first + n;
    *n_rows = (int)n;
    if (n_lost) *n_lost = lost;
    return SG_OK;
}

int time_step_size(Ctx* c, double cfl, double vint, double uref, double vs, double asy, double* dt) {
    StepIn a{c->ni, c->nj, c->udx, c->vdy, (const double*)c->field[SG_F_U], (const double*)c->field[SG_F_V],
             cfl, c->prm.Fr, vint, uref, vs, asy};
    int nb = (c->ni * c->nj + NTS - 1) / NTS;
    if (nb > c->red.max_blocks) nb = c->red.max_blocks;
    if (nb > 2048) nb = 2048;
    double* part = c->red.partials + (size_t)7 * NRED * c->red.max_blocks;
    hipLaunchKernelGGL(k_dt, dim3(nb), dim3(NTS), 0, c->stream, a, part);
    hipLaunchKernelGGL(k_dt_fin, dim3(1), dim3(NTS), 0, c->stream, (const double*)part, nb, c->step_dev);
    SG_HIP(hipGetLastError());
    SG_HIP(hipMemcpyAsync(dt, c->step_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SG_HIP(hipStreamSynchronize(c->stream));
    return SG_OK;
}

// ---- the CFL step without a context and without allocation ---------------
// The per-block partial minima live in a file-scope device array instead of
// the context's reduction buffer.  Host code never passes the array's name
// to a kernel (in host code it names a host-side stand-in): the device
// address is looked up once per device with hipGetSymbolAddress and kept.
constexpr int NOALLOC_PARTIALS = 2048;  // blocks of k_dt, 16 KiB
static __device__ __align__(256) double scratch[NOALLOC_PARTIALS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[MAX_DEVICES] = {nullptr};

int scratch_for_current_device(double** out) {
    int dev = -1;
    SG_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("sg_time_step_size_noalloc: device index beyond the library's table of 64");
        return SG_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error("sg_time_step_size_noalloc: no device address for the scratch array");
            return SG_ERR_HIP;
        }
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return SG_OK;
}
}  // namespace

int time_step_size_noalloc(int ni, int nj, const double* udx, const double* vdy, const double* u, const double* v,
                           double cfl, double Fr, double vint, double uref, double vs, double asy, double* dt_dev,
                           hipStream_t stream) {
    // grid-stride loops cover any grid; the bound keeps e + stride in int
    const long long n = (long long)ni * (long long)nj;
    if (ni < 1 || nj < 1 || n > (long long)INT32_MAX - (long long)NTS * NOALLOC_PARTIALS) {
        set_error("sg_time_step_size_noalloc: ni*nj must be in [1, 2^31 - 1 - 256*2048]");
        return SG_ERR_ARG;
    }
    double* part = nullptr;
    const int rc = scratch_for_current_device(&part);
    if (rc) return rc;
    const StepIn a{ni, nj, udx, vdy, u, v, cfl, Fr, vint, uref, vs, asy};
    long long nb = (n + NTS - 1) / NTS;
    if (nb > NOALLOC_PARTIALS) nb = NOALLOC_PARTIALS;
    // every partial k_dt_fin reads is written by k_dt first: nothing to clear
    hipLaunchKernelGGL(k_dt, dim3((unsigned)nb), dim3(NTS), 0, stream, a, part);
    hipLaunchKernelGGL(k_dt_fin, dim3(1), dim3(NTS), 0, stream, (const double*)part, (int)nb, dt_dev);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int closed_box_bc(Ctx* c) {
    double* u = (double*)c->field[SG_F_U];
    double* v = (double*)c->field[SG_F_V];
    double* p = (double*)c->field[SG_F_P];
    // columns first, then rows (the rows' copies see the corrected columns)
    hipLaunchKernelGGL(k_box_bc, dim3((c->nj + NTS - 1) / NTS), dim3(NTS), 0, c->stream, c->ni, c->nj, u, v, p);
    hipLaunchKernelGGL(k_box_bc_rows, dim3((c->ni + NTS - 1) / NTS), dim3(NTS), 0, c->stream, c->ni, c->nj, u, v, p);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace sg
