// Device helpers shared by the multigrid set-up and cycle kernels.
#pragma once
#include "sg_internal.hpp"

namespace sg {

template <typename T>
struct OpV {
    const T *diag, *cE, *cN, *cNE, *cNW;  // cNE/cNW null on the 5-point level
};

template <typename T>
inline OpV<T> opv(const Op<T>& o) {
    return OpV<T>{o.diag, o.cE, o.cN, o.cNE, o.cNW};
}

template <typename T>
__device__ inline T opget(const T* a, size_t p) {
    return a ? a[p] : T(0);
}

// prolongation weight P(g, D) for fine cell g and coarse point D
template <typename T>
__device__ inline T pweight(const Dims& C, T* const* w, int gi, int gj, int DI, int DJ) {
    const int pi = gi & 1, pj = gj & 1;
    if (!pi && !pj) {
        return (gi >> 1) == DI && (gj >> 1) == DJ ? w[W_INJ][C.off(DI, DJ)] : T(0);
    }
    if (pi && !pj) {
        const int I0 = (gi - 1) >> 1, J0 = gj >> 1;
        if (DJ != J0) return T(0);
        if (DI == I0) return w[W_XW][C.off(I0, J0)];
        if (DI == I0 + 1) return w[W_XE][C.off(I0, J0)];
        return T(0);
    }
    if (!pi && pj) {
        const int I0 = gi >> 1, J0 = (gj - 1) >> 1;
        if (DI != I0) return T(0);
        if (DJ == J0) return w[W_YS][C.off(I0, J0)];
        if (DJ == J0 + 1) return w[W_YN][C.off(I0, J0)];
        return T(0);
    }
    const int I0 = (gi - 1) >> 1, J0 = (gj - 1) >> 1;
    const int a = DI - I0, b = DJ - J0;
    if (a < 0 || a > 1 || b < 0 || b > 1) return T(0);
    const int k = W_SW + a + 2 * b;  // SW, SE, NW, NE
    return w[k][C.off(I0, J0)];
}

// coupling A(f, f+(di,dj)) as a matrix entry (negative of the stored c)
template <typename T>
__device__ inline T aentry(const Dims& F, const OpV<T>& A, int i, int j, int di, int dj) {
    const size_t p = F.off(i, j);
    const size_t ld = F.ld;
    if (di == 0 && dj == 0) return A.diag[p];
    if (dj == 0) return di > 0 ? -A.cE[p] : -A.cE[p - 1];
    if (di == 0) return dj > 0 ? -A.cN[p] : -A.cN[p - ld];
    if (!A.cNE) return T(0);
    if (di > 0 && dj > 0) return -A.cNE[p];
    if (di < 0 && dj < 0) return -A.cNE[p - ld - 1];
    if (di < 0 && dj > 0) return -A.cNW[p];
    return -A.cNW[p - ld + 1];  // (di > 0 && dj < 0)
}

// sum over the neighbours of c * x(neighbour); X(di,dj) reads x
template <typename T, typename XF>
__device__ inline T coupled_sum(const OpV<T>& A, size_t p, size_t ld, XF X) {
    T s = A.cE[p - 1] * X(-1, 0) + A.cE[p] * X(1, 0) + A.cN[p - ld] * X(0, -1) + A.cN[p] * X(0, 1);
    if (A.cNE) {
        s += A.cNE[p] * X(1, 1) + A.cNE[p - ld - 1] * X(-1, -1) + A.cNW[p] * X(-1, 1) + A.cNW[p - ld + 1] * X(1, -1);
    }
    return s;
}

// prolongation of the coarse correction e onto fine cell (gi,gj)
template <typename T>
__device__ inline T prolong_point(const Dims& C, const Prol<T>& P, const T* __restrict__ e, int gi, int gj) {
    T* const* w = P.w;
    const int I = gi >> 1, J = gj >> 1;
    const size_t c = C.off(I, J);
    const int pi = gi & 1, pj = gj & 1;
    if (!pi && !pj) return w[W_INJ][c] * e[c];
    if (pi && !pj) return w[W_XW][c] * e[c] + w[W_XE][c] * e[c + 1];
    if (!pi && pj) return w[W_YS][c] * e[c] + w[W_YN][c] * e[c + C.ld];
    return w[W_SW][c] * e[c] + w[W_SE][c] * e[c + 1] + w[W_NW][c] * e[c + C.ld] + w[W_NE][c] * e[c + C.ld + 1];
}

// restriction P^T r at coarse point (I,J); R(a,b) = r at fine (2I+a, 2J+b)
template <typename T, typename RF>
__device__ inline T restrict_point(const Dims& C, const Prol<T>& P, int I, int J, RF R) {
    T* const* w = P.w;
    const size_t c = C.off(I, J), cw = C.off(I - 1, J), cs = C.off(I, J - 1), csw = C.off(I - 1, J - 1);
    T s = w[W_INJ][c] * R(0, 0);
    s += w[W_XW][c] * R(1, 0) + w[W_XE][cw] * R(-1, 0);
    s += w[W_YS][c] * R(0, 1) + w[W_YN][cs] * R(0, -1);
    s += w[W_SW][c] * R(1, 1) + w[W_SE][cw] * R(-1, 1) + w[W_NW][cs] * R(1, -1) + w[W_NE][csw] * R(-1, -1);
    return s;
}

}  // namespace sg
