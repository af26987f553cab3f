/*
 * CPU FP64 restatement of the projection solver: PCG preconditioned by the
 * structured multigrid V-cycle the GPU path runs -- coarse points on every
 * other cell, operator-dependent (black-box) prolongation, Galerkin 9-point
 * coarse operators P^T A P, red-black Gauss-Seidel on the 5-point fine
 * level, multicolour Gauss-Seidel on the coarse levels, dense solve on the
 * coarsest level.  It stands where HYPRE ParCSR PCG + BoomerAMG stands in the
 * reference (ProjectionP_Mod.f90:52-58,101-135): the converged pressure is
 * fixed by A and b, not by the solver's iterates.
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

#define COARSEST_MAX 16

/* colours of the coarse-level Gauss-Seidel: 4 ((i&1)+2(j&1)) or 2 (i+j&1) */
static int g_coarse_colours = 4;
void orc_set_coarse_colours(int n) { g_coarse_colours = (n == 2) ? 2 : 4; }

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

/* 9-point symmetric operator stored as positive couplings (entry = -c) */
typedef struct {
    int ni, nj;
    double *diag, *cE, *cN, *cNE, *cNW;
    double *x, *b, *r;
    /* prolongation from the next level, at its resolution */
    double* w[9];
} level_t;

enum { W_INJ = 0, W_XW, W_XE, W_YS, W_YN, W_SW, W_SE, W_NW, W_NE };

typedef struct {
    int nl;
    level_t lv[32];
    double* cinv;
    int cn;
} hier_t;

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

#define K(L, i, j) ((size_t)(i) + (size_t)(j) * (size_t)(L)->ni)
static int inb(const level_t* L, int i, int j) { return i >= 0 && j >= 0 && i < L->ni && j < L->nj; }
static double get(const level_t* L, const double* a, int i, int j) { return (a && inb(L, i, j)) ? a[K(L, i, j)] : 0.0; }

/* matrix entry A((i,j),(i+di,j+dj)) */
static double aent(const level_t* L, int i, int j, int di, int dj) {
    if (!inb(L, i, j) || !inb(L, i + di, j + dj)) return 0.0;
    if (di == 0 && dj == 0) return L->diag[K(L, i, j)];
    if (dj == 0) return di > 0 ? -get(L, L->cE, i, j) : -get(L, L->cE, i - 1, j);
    if (di == 0) return dj > 0 ? -get(L, L->cN, i, j) : -get(L, L->cN, i, j - 1);
    if (!L->cNE) return 0.0;
    if (di > 0 && dj > 0) return -get(L, L->cNE, i, j);
    if (di < 0 && dj < 0) return -get(L, L->cNE, i - 1, j - 1);
    if (di < 0 && dj > 0) return -get(L, L->cNW, i, j);
    return -get(L, L->cNW, i + 1, j - 1);
}

/* coupling of (i,j) to (i+di,j+dj) as a positive number */
static double cpl(const level_t* L, int i, int j, int di, int dj) { return -aent(L, i, j, di, dj); }

/* P(g, D): weight of coarse point D on fine cell g; C = coarse level holding w */
static double pw(const level_t* F, const level_t* C, int gi, int gj, int DI, int DJ) {
    if (gi < 0 || gj < 0 || gi >= F->ni || gj >= F->nj) return 0.0;
    const int pi = gi & 1, pj = gj & 1;
    double* const* w = F->w;
    if (!pi && !pj) return (gi / 2 == DI && gj / 2 == DJ) ? get(C, w[W_INJ], DI, DJ) : 0.0;
    if (pi && !pj) {
        const int I0 = (gi - 1) / 2, J0 = gj / 2;
        if (DJ != J0) return 0.0;
        if (DI == I0) return get(C, w[W_XW], I0, J0);
        if (DI == I0 + 1) return get(C, w[W_XE], I0, J0);
        return 0.0;
    }
    if (!pi && pj) {
        const int I0 = gi / 2, J0 = (gj - 1) / 2;
        if (DI != I0) return 0.0;
        if (DJ == J0) return get(C, w[W_YS], I0, J0);
        if (DJ == J0 + 1) return get(C, w[W_YN], I0, J0);
        return 0.0;
    }
    const int I0 = (gi - 1) / 2, J0 = (gj - 1) / 2;
    const int a = DI - I0, b = DJ - J0;
    if (a < 0 || a > 1 || b < 0 || b > 1) return 0.0;
    return get(C, w[W_SW + a + 2 * b], I0, J0);
}

/* black-box prolongation weights of fine level F, stored on C's grid */
static void prolongation(level_t* F, const level_t* C) {
    for (int k = 0; k < 9; ++k) F->w[k] = dalloc((size_t)C->ni * C->nj);
#pragma omp parallel for
    for (int J = 0; J < C->nj; ++J)
        for (int I = 0; I < C->ni; ++I) {
            const size_t c = K(C, I, J);
            const int fi = 2 * I, fj = 2 * J;
            F->w[W_INJ][c] = F->diag[K(F, fi, fj)] > 0.0 ? 1.0 : 0.0;
            if (fi + 1 < F->ni) {
                const int i = fi + 1, j = fj;
                const double d = F->diag[K(F, i, j)];
                if (d > 0.0) {
                    const double den = d - cpl(F, i, j, 0, 1) - cpl(F, i, j, 0, -1);
                    const double wc = cpl(F, i, j, -1, 0) + cpl(F, i, j, -1, 1) + cpl(F, i, j, -1, -1);
                    const double ec = cpl(F, i, j, 1, 0) + cpl(F, i, j, 1, 1) + cpl(F, i, j, 1, -1);
                    if (den > 0.0) {
                        F->w[W_XW][c] = wc / den;
                        F->w[W_XE][c] = ec / den;
                    }
                }
            }
            if (fj + 1 < F->nj) {
                const int i = fi, j = fj + 1;
                const double d = F->diag[K(F, i, j)];
                if (d > 0.0) {
                    const double den = d - cpl(F, i, j, 1, 0) - cpl(F, i, j, -1, 0);
                    const double sr = cpl(F, i, j, 0, -1) + cpl(F, i, j, -1, -1) + cpl(F, i, j, 1, -1);
                    const double nr = cpl(F, i, j, 0, 1) + cpl(F, i, j, -1, 1) + cpl(F, i, j, 1, 1);
                    if (den > 0.0) {
                        F->w[W_YS][c] = sr / den;
                        F->w[W_YN][c] = nr / den;
                    }
                }
            }
        }
#pragma omp parallel for
    for (int J = 0; J < C->nj; ++J)
        for (int I = 0; I < C->ni; ++I) {
            const size_t c = K(C, I, J);
            const int i = 2 * I + 1, j = 2 * J + 1;
            if (i >= F->ni || j >= F->nj) continue;
            const double d = F->diag[K(F, i, j)];
            if (!(d > 0.0)) continue;
            const double cW = cpl(F, i, j, -1, 0), cE = cpl(F, i, j, 1, 0);
            const double cS = cpl(F, i, j, 0, -1), cN = cpl(F, i, j, 0, 1);
            const double cSW = cpl(F, i, j, -1, -1), cSE = cpl(F, i, j, 1, -1);
            const double cNW = cpl(F, i, j, -1, 1), cNE = cpl(F, i, j, 1, 1);
            double* const* w = F->w;
#define WG(k, a, b) get(C, w[k], a, b)
            F->w[W_SW][c] = (cW * WG(W_YS, I, J) + cS * WG(W_XW, I, J) + cSW * WG(W_INJ, I, J)) / d;
            F->w[W_SE][c] = (cE * WG(W_YS, I + 1, J) + cS * WG(W_XE, I, J) + cSE * WG(W_INJ, I + 1, J)) / d;
            F->w[W_NW][c] = (cW * WG(W_YN, I, J) + cN * WG(W_XW, I, J + 1) + cNW * WG(W_INJ, I, J + 1)) / d;
            F->w[W_NE][c] = (cE * WG(W_YN, I + 1, J) + cN * WG(W_XE, I, J + 1) + cNE * WG(W_INJ, I + 1, J + 1)) / d;
#undef WG
        }
}

/* C->op = P^T F->op P */
static void galerkin(const level_t* F, level_t* C) {
#pragma omp parallel for
    for (int J = 0; J < C->nj; ++J)
        for (int I = 0; I < C->ni; ++I) {
            const int DI[5] = {I, I + 1, I, I + 1, I - 1};
            const int DJ[5] = {J, J, J + 1, J + 1, J + 1};
            double acc[5] = {0, 0, 0, 0, 0};
            for (int b = -1; b <= 1; ++b)
                for (int a = -1; a <= 1; ++a) {
                    const int fi = 2 * I + a, fj = 2 * J + b;
                    const double pc = pw(F, C, fi, fj, I, J);
                    if (pc == 0.0) continue;
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            const double afg = aent(F, fi, fj, di, dj);
                            if (afg == 0.0) continue;
                            for (int k = 0; k < 5; ++k) acc[k] += pc * afg * pw(F, C, fi + di, fj + dj, DI[k], DJ[k]);
                        }
                }
            const size_t c = K(C, I, J);
            C->diag[c] = acc[0] > 0.0 ? acc[0] : 0.0;
            C->cE[c] = (I + 1 < C->ni) ? -acc[1] : 0.0;
            C->cN[c] = (J + 1 < C->nj) ? -acc[2] : 0.0;
            C->cNE[c] = (I + 1 < C->ni && J + 1 < C->nj) ? -acc[3] : 0.0;
            C->cNW[c] = (I > 0 && J + 1 < C->nj) ? -acc[4] : 0.0;
        }
}

static void coarse_inverse(const level_t* c, double* inv) {
    const int n = c->ni * c->nj;
    double* a = dalloc((size_t)n * n);
    for (int r = 0; r < n; ++r)
        for (int q = 0; q < n; ++q) {
            const int ri = r % c->ni, rj = r / c->ni, qi = q % c->ni, qj = q / c->ni;
            const int di = qi - ri, dj = qj - rj;
            a[(size_t)r * n + q] = (abs(di) <= 1 && abs(dj) <= 1) ? aent(c, ri, rj, di, dj) : 0.0;
            inv[(size_t)r * n + q] = r == q ? 1.0 : 0.0;
        }
    double amax = 0.0;
    for (int k = 0; k < n; ++k) amax = fmax(amax, fabs(a[(size_t)k * n + k]));
    for (int k = 0; k < n; ++k) {
        const double piv = a[(size_t)k * n + k];
        if (!(fabs(piv) > 1e-13 * amax)) {
            for (int m = 0; m < n; ++m) {
                a[(size_t)k * n + m] = a[(size_t)m * n + k] = 0.0;
                inv[(size_t)k * n + m] = inv[(size_t)m * n + k] = 0.0;
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

static void alloc_level(level_t* L, int ni, int nj, int nine) {
    memset(L, 0, sizeof(*L));
    L->ni = ni;
    L->nj = nj;
    const size_t n = (size_t)ni * nj;
    L->diag = dalloc(n);
    L->cE = dalloc(n);
    L->cN = dalloc(n);
    if (nine) {
        L->cNE = dalloc(n);
        L->cNW = dalloc(n);
    }
    L->x = dalloc(n);
    L->b = dalloc(n);
    L->r = dalloc(n);
}

/* A fluid row whose faces are all closed keeps only the regularisation
 * D = +1e-24 of ProjectionP_Mod.f90:227 and no coupling: it is a 1x1 system
 * of its own, solved exactly as b / D after the PCG; the hierarchy treats it
 * as a solid row. */
static int decoupled(const double* matr, size_t k, size_t N) {
    return matr[k] == 0.0 && matr[k + N] == 0.0 && matr[k + 3 * N] == 0.0 && matr[k + 4 * N] == 0.0 &&
           matr[k + 2 * N] != 0.0 && fabs(matr[k + 2 * N]) <= 1.0e-24;
}

static void build_hier(hier_t* h, int ni, int nj, const double* matr) {
    memset(h, 0, sizeof(*h));
    const size_t N = (size_t)ni * nj;
    level_t* f = &h->lv[0];
    alloc_level(f, ni, nj, 0);
    for (size_t k = 0; k < N; ++k) {
        f->diag[k] = (matr[k + 2 * N] > 0.0 && !decoupled(matr, k, N)) ? matr[k + 2 * N] : 0.0;
        f->cE[k] = -matr[k + 3 * N];
        f->cN[k] = -matr[k + 4 * N];
    }
    int l = 0;
    while ((size_t)h->lv[l].ni * h->lv[l].nj > COARSEST_MAX && l < 30) {
        level_t* c = &h->lv[l + 1];
        alloc_level(c, (h->lv[l].ni + 1) / 2, (h->lv[l].nj + 1) / 2, 1);
        prolongation(&h->lv[l], c);
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
        free(v->diag); free(v->cE); free(v->cN); free(v->cNE); free(v->cNW);
        free(v->x); free(v->b); free(v->r);
        for (int k = 0; k < 9; ++k) free(v->w[k]);
    }
    free(h->cinv);
}

/* one Gauss-Seidel pass over the points of one colour (Jacobi within the
 * colour for couplings between points of the same colour) */
static void gs_colour(level_t* L, int ncol, int col) {
    const int ni = L->ni, nj = L->nj;
    double* xn = L->r; /* new values of this colour, then copied */
#pragma omp parallel for
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const int c = ncol == 2 ? ((i + j) & 1) : ((i & 1) + 2 * (j & 1));
            if (c != col) continue;
            const size_t k = K(L, i, j);
            const double d = L->diag[k];
            if (!(d > 0.0)) { xn[k] = 0.0; continue; }
            double s = L->b[k];
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    if (!di && !dj) continue;
                    const double a = aent(L, i, j, di, dj);
                    if (a != 0.0) s -= a * L->x[K(L, i + di, j + dj)];
                }
            xn[k] = s / d;
        }
#pragma omp parallel for
    for (int j = 0; j < nj; ++j)
        for (int i = 0; i < ni; ++i) {
            const int c = ncol == 2 ? ((i + j) & 1) : ((i & 1) + 2 * (j & 1));
            if (c == col) L->x[K(L, i, j)] = xn[K(L, i, j)];
        }
}

static void smooth(level_t* L, int level, int sweeps, int reverse) {
    const int ncol = (level == 0) ? 2 : g_coarse_colours;
    for (int s = 0; s < sweeps; ++s)
        for (int k = 0; k < ncol; ++k) gs_colour(L, ncol, reverse ? ncol - 1 - k : k);
}

static void lv_apply(const level_t* L, const double* x, double* y) {
#pragma omp parallel for
    for (int j = 0; j < L->nj; ++j)
        for (int i = 0; i < L->ni; ++i) {
            double s = 0.0;
            for (int dj = -1; dj <= 1; ++dj)
                for (int di = -1; di <= 1; ++di) {
                    const double a = aent(L, i, j, di, dj);
                    if (a != 0.0) s += a * x[K(L, i + di, j + dj)];
                }
            y[K(L, i, j)] = s;
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
    level_t* C = &h->lv[l + 1];
    memset(L->x, 0, N * sizeof(double));
    smooth(L, l, pre, 0);
    lv_apply(L, L->x, L->r);
    for (size_t k = 0; k < N; ++k) L->r[k] = L->b[k] - L->r[k];
    /* restriction P^T r */
#pragma omp parallel for
    for (int J = 0; J < C->nj; ++J)
        for (int I = 0; I < C->ni; ++I) {
            double s = 0.0;
            for (int b = -1; b <= 1; ++b)
                for (int a = -1; a <= 1; ++a) {
                    const int fi = 2 * I + a, fj = 2 * J + b;
                    if (!inb(L, fi, fj)) continue;
                    const double w = pw(L, C, fi, fj, I, J);
                    if (w != 0.0) s += w * L->r[K(L, fi, fj)];
                }
            C->b[K(C, I, J)] = s;
        }
    vcycle(h, l + 1, pre, post);
    /* x += P e */
#pragma omp parallel for
    for (int j = 0; j < L->nj; ++j)
        for (int i = 0; i < L->ni; ++i) {
            double s = 0.0;
            const int I0 = (i - 1) / 2 < 0 ? 0 : (i - 1) / 2;
            for (int J = (j - 1) / 2 < 0 ? 0 : (j - 1) / 2; J <= (j + 1) / 2 && J < C->nj; ++J)
                for (int I = I0; I <= (i + 1) / 2 && I < C->ni; ++I) {
                    const double w = pw(L, C, i, j, I, J);
                    if (w != 0.0) s += w * C->x[K(C, I, J)];
                }
            L->x[K(L, i, j)] += s;
        }
    smooth(L, l, post, 1);
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
        for (size_t k = 0; k < N; ++k)
            if (decoupled(matr, k, N)) x[k] = b[k] / matr[k + 2 * N];
        if (relres_out) *relres_out = 0.0;
        free_hier(&h); free(r); free(z); free(p); free(q);
        return 0;
    }
    double rz = 0.0;
    int first = 1;
    for (int rep = 0; rep < 5; ++rep) {
        while (rn > tol * bn && it < max_iter) {
            memcpy(F->b, r, N * sizeof(double));
            vcycle(&h, 0, pre, post);
            memcpy(z, F->x, N * sizeof(double));
            const double rz_new = dot(r, z, N);
            if (first) {
                memcpy(p, z, N * sizeof(double));
                first = 0;
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
        /* residual replacement: restart from the true residual */
        lv_apply(F, x, q);
        for (size_t k = 0; k < N; ++k) r[k] = (F->diag[k] > 0.0) ? b[k] - q[k] : 0.0;
        rn = sqrt(dot(r, r, N));
        first = 1;
        if (rn <= tol * bn || it >= max_iter) break;
    }
    for (size_t k = 0; k < N; ++k)
        if (decoupled(matr, k, N)) x[k] = b[k] / matr[k + 2 * N];
    if (relres_out) *relres_out = rn / bn;
    free_hier(&h); free(r); free(z); free(p); free(q);
    return (rn > tol * bn) ? -it : it;
}

int orc_pcg_jacobi_solve(int ni, int nj, const double* matr, const double* b,
                         double* x, double tol, int max_iter, double* relres_out) {
    level_t F;
    const size_t N = (size_t)ni * nj;
    alloc_level(&F, ni, nj, 0);
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
    free(F.diag); free(F.cE); free(F.cN); free(F.x); free(F.b); free(F.r);
    free(r); free(z); free(p); free(q);
    return (bn > 0.0 && rn > tol * bn) ? -it : it;
}
