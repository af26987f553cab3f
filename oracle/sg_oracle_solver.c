/*
 * CPU FP64 restatement of the projection solver: PCG preconditioned by a
 * structured geometric multigrid V-cycle (2x2 aggregation with Galerkin
 * coarse operators, red-black Gauss-Seidel smoothing, direct solve on the
 * coarsest level).  It stands where HYPRE ParCSR PCG + BoomerAMG stands in
 * the reference (ProjectionP_Mod.f90:52-58,101-135): the converged pressure
 * is fixed by A and b, not by the solver's iterates.
 *
 * TEST INFRASTRUCTURE ONLY: the checker and the timed CPU baseline.
 */
#include "sg_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define COARSEST_MAX 64

static double g_cscale = 1.0;
void orc_set_coarse_scale(double s) { g_cscale = s; }
static int g_gamma = 1;
void orc_set_cycle(int g) { g_gamma = g; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

typedef struct {
    int ni, nj;
    double *cE, *cN, *diag, *x, *b, *r;
} level_t;

typedef struct {
    int nl;
    level_t lv[32];
    double* cinv; /* dense inverse of the coarsest operator, n x n */
    int cn;
} hier_t;

#define IX(l, i, j) ((size_t)(i) + (size_t)(j) * (size_t)(l)->ni)

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

/* Galerkin coarse operator for piecewise-constant prolongation over 2x2
 * aggregates (solid cells have diag == 0 and zero faces, so they drop out). */
static void galerkin(const level_t* f, level_t* c) {
#pragma omp parallel for
    for (int J = 0; J < c->nj; ++J) {
        for (int I = 0; I < c->ni; ++I) {
            double d = 0.0, e = 0.0, n = 0.0;
            const int i0 = 2 * I, j0 = 2 * J;
            const int i1 = i0 + 1 < f->ni ? i0 + 1 : -1;
            const int j1 = j0 + 1 < f->nj ? j0 + 1 : -1;
            d += f->diag[IX(f, i0, j0)];
            if (i1 >= 0) d += f->diag[IX(f, i1, j0)];
            if (j1 >= 0) d += f->diag[IX(f, i0, j1)];
            if (i1 >= 0 && j1 >= 0) d += f->diag[IX(f, i1, j1)];
            if (i1 >= 0) {
                d -= 2.0 * f->cE[IX(f, i0, j0)];
                if (j1 >= 0) d -= 2.0 * f->cE[IX(f, i0, j1)];
            }
            if (j1 >= 0) {
                d -= 2.0 * f->cN[IX(f, i0, j0)];
                if (i1 >= 0) d -= 2.0 * f->cN[IX(f, i1, j0)];
            }
            const int ie = i1 >= 0 ? i1 : i0;
            e += f->cE[IX(f, ie, j0)];
            if (j1 >= 0) e += f->cE[IX(f, ie, j1)];
            const int jn = j1 >= 0 ? j1 : j0;
            n += f->cN[IX(f, i0, jn)];
            if (i1 >= 0) n += f->cN[IX(f, i1, jn)];
            if (I == c->ni - 1) e = 0.0;
            if (J == c->nj - 1) n = 0.0;
            c->diag[IX(c, I, J)] = d > 0.0 ? g_cscale * d : 0.0;
            c->cE[IX(c, I, J)] = g_cscale * e;
            c->cN[IX(c, I, J)] = g_cscale * n;
        }
    }
}

/* Dense inverse of the coarsest operator by Gauss-Jordan with a zero-pivot
 * guard (a zero pivot -- a solid cell, or the constant mode of an all-Neumann
 * problem -- maps to a zero row/column). */
static void coarse_inverse(const level_t* c, double* inv) {
    const int n = c->ni * c->nj;
    double* a = dalloc((size_t)n * n);
    memset(inv, 0, sizeof(double) * (size_t)n * n);
    for (int j = 0; j < c->nj; ++j)
        for (int i = 0; i < c->ni; ++i) {
            const int k = i + j * c->ni;
            a[(size_t)k * n + k] = c->diag[k];
            if (i + 1 < c->ni) {
                a[(size_t)k * n + k + 1] = -c->cE[k];
                a[(size_t)(k + 1) * n + k] = -c->cE[k];
            }
            if (j + 1 < c->nj) {
                a[(size_t)k * n + k + c->ni] = -c->cN[k];
                a[(size_t)(k + c->ni) * n + k] = -c->cN[k];
            }
        }
    for (int k = 0; k < n; ++k) inv[(size_t)k * n + k] = 1.0;
    double amax = 0.0;
    for (int k = 0; k < n; ++k) amax = fmax(amax, fabs(a[(size_t)k * n + k]));
    for (int k = 0; k < n; ++k) {
        const double piv = a[(size_t)k * n + k];
        if (!(fabs(piv) > 1e-13 * amax)) {
            for (int m = 0; m < n; ++m) {
                a[(size_t)k * n + m] = 0.0;
                a[(size_t)m * n + k] = 0.0;
                inv[(size_t)k * n + m] = 0.0;
                inv[(size_t)m * n + k] = 0.0;
            }
            continue;
        }
        const double ip = 1.0 / piv;
        for (int m = 0; m < n; ++m) {
            a[(size_t)k * n + m] *= ip;
            inv[(size_t)k * n + m] *= ip;
        }
        for (int r = 0; r < n; ++r) {
            if (r == k) continue;
            const double f = a[(size_t)r * n + k];
            if (f == 0.0) continue;
            for (int m = 0; m < n; ++m) {
                a[(size_t)r * n + m] -= f * a[(size_t)k * n + m];
                inv[(size_t)r * n + m] -= f * inv[(size_t)k * n + m];
            }
        }
    }
    free(a);
}

static void build_hier(hier_t* h, int ni, int nj, const double* matr) {
    memset(h, 0, sizeof(*h));
    const size_t N = (size_t)ni * nj;
    level_t* f = &h->lv[0];
    f->ni = ni;
    f->nj = nj;
    f->cE = dalloc(N);
    f->cN = dalloc(N);
    f->diag = dalloc(N);
    f->x = dalloc(N);
    f->b = dalloc(N);
    f->r = dalloc(N);
    /* symmetric form: the E/N entry of each row is the face coefficient */
    for (size_t k = 0; k < N; ++k) {
        f->diag[k] = matr[k + 2 * N] > 0.0 ? matr[k + 2 * N] : 0.0;
        f->cE[k] = -matr[k + 3 * N];
        f->cN[k] = -matr[k + 4 * N];
    }
    int l = 0;
    while ((size_t)h->lv[l].ni * h->lv[l].nj > COARSEST_MAX && l < 30) {
        level_t* c = &h->lv[l + 1];
        c->ni = (h->lv[l].ni + 1) / 2;
        c->nj = (h->lv[l].nj + 1) / 2;
        const size_t M = (size_t)c->ni * c->nj;
        c->cE = dalloc(M);
        c->cN = dalloc(M);
        c->diag = dalloc(M);
        c->x = dalloc(M);
        c->b = dalloc(M);
        c->r = dalloc(M);
        galerkin(&h->lv[l], c);
        ++l;
    }
    h->nl = l + 1;
    h->cn = h->lv[l].ni * h->lv[l].nj;
    h->cinv = dalloc((size_t)h->cn * h->cn);
    coarse_inverse(&h->lv[l], h->cinv);
}

static void free_hier(hier_t* h) {
    for (int l = 0; l < h->nl; ++l) {
        level_t* v = &h->lv[l];
        free(v->cE); free(v->cN); free(v->diag); free(v->x); free(v->b); free(v->r);
    }
    free(h->cinv);
}

/* one half-sweep of red-black Gauss-Seidel on colour col ((i+j)&1 == col) */
static void rb_half(level_t* L, int col) {
    const int ni = L->ni, nj = L->nj;
#pragma omp parallel for
    for (int j = 0; j < nj; ++j) {
        for (int i = ((j + col) & 1); i < ni; i += 2) {
            const size_t k = IX(L, i, j);
            const double d = L->diag[k];
            if (!(d > 0.0)) { L->x[k] = 0.0; continue; }
            double s = L->b[k];
            if (i > 0) s += L->cE[k - 1] * L->x[k - 1];
            if (i + 1 < ni) s += L->cE[k] * L->x[k + 1];
            if (j > 0) s += L->cN[k - ni] * L->x[k - ni];
            if (j + 1 < nj) s += L->cN[k] * L->x[k + ni];
            L->x[k] = s / d;
        }
    }
}

/* y = A x for a level operator */
static void lv_apply(const level_t* L, const double* x, double* y) {
    const int ni = L->ni, nj = L->nj;
#pragma omp parallel for
    for (int j = 0; j < nj; ++j) {
        for (int i = 0; i < ni; ++i) {
            const size_t k = IX(L, i, j);
            double s = L->diag[k] * x[k];
            if (i > 0) s -= L->cE[k - 1] * x[k - 1];
            if (i + 1 < ni) s -= L->cE[k] * x[k + 1];
            if (j > 0) s -= L->cN[k - ni] * x[k - ni];
            if (j + 1 < nj) s -= L->cN[k] * x[k + ni];
            y[k] = s;
        }
    }
}

static void vcycle(hier_t* h, int l, int pre, int post) {
    level_t* L = &h->lv[l];
    const size_t N = (size_t)L->ni * L->nj;
    if (l == h->nl - 1) {
        const int n = h->cn;
        for (int r = 0; r < n; ++r) {
            double s = 0.0;
            for (int m = 0; m < n; ++m) s += h->cinv[(size_t)r * n + m] * L->b[m];
            L->x[r] = s;
        }
        return;
    }
    memset(L->x, 0, N * sizeof(double));
    for (int s = 0; s < pre; ++s) { rb_half(L, 0); rb_half(L, 1); }
    lv_apply(L, L->x, L->r);
    level_t* c = &h->lv[l + 1];
#pragma omp parallel for
    for (int J = 0; J < c->nj; ++J)
        for (int I = 0; I < c->ni; ++I) {
            double s = 0.0;
            for (int b = 0; b < 2; ++b)
                for (int a = 0; a < 2; ++a) {
                    const int i = 2 * I + a, j = 2 * J + b;
                    if (i < L->ni && j < L->nj) {
                        const size_t k = IX(L, i, j);
                        s += L->b[k] - L->r[k];
                    }
                }
            c->b[IX(c, I, J)] = s;
        }
    vcycle(h, l + 1, pre, post);
    for (int gg = 1; gg < g_gamma && l + 1 < h->nl - 1; ++gg) {
        /* extra coarse-grid iterations: residual of the coarse problem */
        level_t* cc = &h->lv[l + 1];
        const size_t M = (size_t)cc->ni * cc->nj;
        double* save_b = dalloc(M); double* save_x = dalloc(M);
        memcpy(save_b, cc->b, M * sizeof(double)); memcpy(save_x, cc->x, M * sizeof(double));
        lv_apply(cc, cc->x, cc->r);
        for (size_t k = 0; k < M; ++k) cc->b[k] = save_b[k] - cc->r[k];
        vcycle(h, l + 1, pre, post);
        for (size_t k = 0; k < M; ++k) cc->x[k] += save_x[k];
        memcpy(cc->b, save_b, M * sizeof(double));
        free(save_b); free(save_x);
    }
#pragma omp parallel for
    for (int j = 0; j < L->nj; ++j)
        for (int i = 0; i < L->ni; ++i) {
            const size_t k = IX(L, i, j);
            if (L->diag[k] > 0.0) L->x[k] += c->x[IX(c, i / 2, j / 2)];
        }
    for (int s = 0; s < post; ++s) { rb_half(L, 1); rb_half(L, 0); }
}

static double dot(const double* a, const double* b, size_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s)
    for (size_t k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

int orc_pcg_mg_solve(int ni, int nj, const double* matr, const double* b,
                     double* x, double tol, int max_iter, int pre, int post,
                     double* relres_out) {
    hier_t h;
    build_hier(&h, ni, nj, matr);
    level_t* F = &h.lv[0];
    const size_t N = (size_t)ni * nj;
    double *r = dalloc(N), *z = dalloc(N), *p = dalloc(N), *q = dalloc(N);
    /* solid rows carry no unknown */
    for (size_t k = 0; k < N; ++k) if (!(F->diag[k] > 0.0)) x[k] = 0.0;
    lv_apply(F, x, q);
    for (size_t k = 0; k < N; ++k) r[k] = (F->diag[k] > 0.0) ? b[k] - q[k] : 0.0;
    double bn = 0.0;
    for (size_t k = 0; k < N; ++k) if (F->diag[k] > 0.0) bn += b[k] * b[k];
    bn = sqrt(bn);
    double rn = sqrt(dot(r, r, N));
    int it = 0;
    if (bn == 0.0) {
        memset(x, 0, N * sizeof(double));
        if (relres_out) *relres_out = 0.0;
        free_hier(&h); free(r); free(z); free(p); free(q);
        return 0;
    }
    double rz = 0.0;
    while (rn > tol * bn && it < max_iter) {
        memcpy(F->b, r, N * sizeof(double));
        vcycle(&h, 0, pre, post);
        memcpy(z, F->x, N * sizeof(double));
        const double rz_new = dot(r, z, N);
        if (it == 0) {
            memcpy(p, z, N * sizeof(double));
        } else {
            const double beta = rz_new / rz;
            for (size_t k = 0; k < N; ++k) p[k] = z[k] + beta * p[k];
        }
        rz = rz_new;
        lv_apply(F, p, q);
        const double alpha = rz / dot(p, q, N);
        for (size_t k = 0; k < N; ++k) {
            x[k] += alpha * p[k];
            r[k] -= alpha * q[k];
        }
        rn = sqrt(dot(r, r, N));
        ++it;
    }
    /* report the true residual */
    lv_apply(F, x, q);
    double tr = 0.0;
    for (size_t k = 0; k < N; ++k) if (F->diag[k] > 0.0) tr += (b[k] - q[k]) * (b[k] - q[k]);
    if (relres_out) *relres_out = sqrt(tr) / bn;
    free_hier(&h); free(r); free(z); free(p); free(q);
    return (rn > tol * bn) ? -it : it;
}

int orc_pcg_jacobi_solve(int ni, int nj, const double* matr, const double* b,
                         double* x, double tol, int max_iter, double* relres_out) {
    level_t F;
    const size_t N = (size_t)ni * nj;
    F.ni = ni; F.nj = nj;
    F.cE = dalloc(N); F.cN = dalloc(N); F.diag = dalloc(N);
    for (size_t k = 0; k < N; ++k) {
        F.diag[k] = matr[k + 2 * N] > 0.0 ? matr[k + 2 * N] : 0.0;
        F.cE[k] = -matr[k + 3 * N];
        F.cN[k] = -matr[k + 4 * N];
    }
    double *r = dalloc(N), *z = dalloc(N), *p = dalloc(N), *q = dalloc(N);
    for (size_t k = 0; k < N; ++k) if (!(F.diag[k] > 0.0)) x[k] = 0.0;
    lv_apply(&F, x, q);
    double bn = 0.0;
    for (size_t k = 0; k < N; ++k) {
        r[k] = (F.diag[k] > 0.0) ? b[k] - q[k] : 0.0;
        if (F.diag[k] > 0.0) bn += b[k] * b[k];
    }
    bn = sqrt(bn);
    double rn = sqrt(dot(r, r, N)), rz = 0.0;
    int it = 0;
    while (bn > 0.0 && rn > tol * bn && it < max_iter) {
        for (size_t k = 0; k < N; ++k) z[k] = F.diag[k] > 0.0 ? r[k] / F.diag[k] : 0.0;
        const double rz_new = dot(r, z, N);
        if (it == 0) memcpy(p, z, N * sizeof(double));
        else {
            const double beta = rz_new / rz;
            for (size_t k = 0; k < N; ++k) p[k] = z[k] + beta * p[k];
        }
        rz = rz_new;
        lv_apply(&F, p, q);
        const double alpha = rz / dot(p, q, N);
        for (size_t k = 0; k < N; ++k) { x[k] += alpha * p[k]; r[k] -= alpha * q[k]; }
        rn = sqrt(dot(r, r, N));
        ++it;
    }
    if (relres_out) *relres_out = bn > 0.0 ? rn / bn : 0.0;
    free(F.cE); free(F.cN); free(F.diag); free(r); free(z); free(p); free(q);
    return (bn > 0.0 && rn > tol * bn) ? -it : it;
}
