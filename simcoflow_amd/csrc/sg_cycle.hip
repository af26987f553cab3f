// Multigrid V-cycle kernels: one fused kernel per level on the way down
// (red-black pre-smoothing from a zero guess + residual + restriction),
// one on the way up (prolongation + red-black post-smoothing), and one
// single-workgroup kernel for all levels of at most TAIL_MAX points.
//
// Every kernel stages its operands (b, x, the operator) for a tile plus the
// halo its passes need in LDS with one unrolled load phase, then runs the
// colour passes out of LDS: a sweep reads each coefficient once from HBM
// and pays one memory latency instead of one per pass.  On the 9-point
// coarse levels the red-black passes treat couplings between cells of the
// same colour (the diagonal ones) Jacobi-wise: new values are staged in
// registers and written after a barrier, which keeps the update independent
// of thread order and the cycle symmetric (post-smoothing runs the colours
// in reverse order).
#include "sg_internal.hpp"
#include "sg_mg_common.hpp"
#include "sg_reduce.hpp"

namespace sg {

namespace {

constexpr int NTB = 256;

constexpr int TAIL_MAX = 1024;   // levels with at most this many points run in the tail kernel
constexpr int TAIL_THREADS = 1024;
constexpr int MAX_TAIL = 10;
constexpr size_t TAIL_LDS_MAX = 150 * 1024;  // of the 160 KB of LDS a workgroup can hold

struct Slots {
    double* part;
    int max_blocks;
};

__device__ inline bool in_dom(const Dims& D, int i, int j) { return i >= 0 && j >= 0 && i < D.ni && j < D.nj; }

// ---------------------------------------------------------------------------
// 9-point coarse levels that are not in the tail: the operator is read from
// global memory (through the caches) and only b, x live in LDS, which keeps
// occupancy high.  Couplings between cells of one colour (the diagonals) are
// taken Jacobi-wise: a colour pass computes into registers and writes after
// a barrier.

// A 9-point row's couplings (d, west, east, south, north, NE, SW, NW, SE),
// loaded ahead of the passes that use them
template <typename T>
struct Row9 {
    T d, w, e, s, n, ne, sw, nw, se;
};
template <typename T>
__device__ __forceinline__ Row9<T> load_row9(const OpV<T>& A, size_t p, size_t ld) {
    Row9<T> r;
    r.d = A.diag[p];
    r.w = A.cE[p - 1];
    r.e = A.cE[p];
    r.s = A.cN[p - ld];
    r.n = A.cN[p];
    if (A.cNE) {
        r.ne = A.cNE[p];
        r.sw = A.cNE[p - ld - 1];
        r.nw = A.cNW[p];
        r.se = A.cNW[p - ld + 1];
    } else {
        r.ne = r.sw = r.nw = r.se = T(0);
    }
    return r;
}
// coupled_sum over a loaded row (same terms, same order)
template <typename T, typename XF>
__device__ __forceinline__ T row9_sum(const Row9<T>& r, bool nine, XF X) {
    T s = r.w * X(-1, 0) + r.e * X(1, 0) + r.s * X(0, -1) + r.n * X(0, 1);
    if (nine) s += r.ne * X(1, 1) + r.sw * X(-1, -1) + r.nw * X(-1, 1) + r.se * X(1, -1);
    return s;
}

// down: x = S(0; b) (red from zero, then black), r = b - A x, bc = P^T r.
// With tiles of at most NTB cells every coupling and restriction weight the
// passes read is loaded before the first barrier (a few cells per thread):
// the passes then wait on LDS only, which matters on the small levels, where
// a kernel is a chain of memory latencies.
template <typename T, int TW, int TH>
__global__ void __launch_bounds__(NTB) k_down9(Dims F, OpV<T> A, const T* __restrict__ b, T* __restrict__ xo,
                                               Dims C, Prol<T> P, T* __restrict__ bc, const int* on) {
    if (on && !*on) return;
    constexpr int RW = TW + 5, RH = TH + 5;  // local (li,lj) <-> cell (i0+li-3, j0+lj-3)
    constexpr int NL = (RW * RH + NTB - 1) / NTB;
    __shared__ T sb[RH][RW];
    __shared__ T sx[RH][RW];
    int tx, ty;
    xcd_tile(tx, ty);
    const int i0 = tx * TW, j0 = ty * TH;
    const int tid = threadIdx.x;
    const size_t ld = F.ld;
#pragma unroll
    for (int m = 0; m < NL; ++m) {
        const int k = tid + m * NTB;
        if (k >= RW * RH) continue;
        const int li = k % RW, lj = k / RW;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        T bv = T(0), xv = T(0);
        if (in_dom(F, gi, gj)) {
            const size_t p = F.off(gi, gj);
            bv = b[p];
            if (((gi + gj) & 1) == 0) {
                const T d = A.diag[p];
                xv = d > T(0) ? bv / d : T(0);
            }
        }
        sb[lj][li] = bv;
        sx[lj][li] = xv;
    }
    constexpr bool PRE = TW * TH <= NTB;
    const bool nine = A.cNE != nullptr;
    constexpr int PR = (RW - 1) / 2;
    constexpr int NBK = (PR * (RH - 2) + NTB - 1) / NTB;
    constexpr int NRS = ((TW + 1) * (TH + 1) + NTB - 1) / NTB;
    // black cell n of this thread: local (li, lj), or li < 0
    auto black_cell = [&](int n, int& li, int& lj) {
        const int k = tid + n * NTB;
        li = -1;
        lj = 0;
        if (k >= PR * (RH - 2)) return;
        lj = k / PR + 1;
        const int l = 2 * (k % PR) + 1 + (lj & 1);
        if (l >= RW - 1 || !in_dom(F, i0 + l - 3, j0 + lj - 3)) return;
        li = l;
    };
    Row9<T> rbk[PRE ? NBK : 1], rrs[PRE ? NRS : 1];
    T rw[9];
    if constexpr (PRE) {
#pragma unroll
        for (int n = 0; n < NBK; ++n) {
            int li, lj;
            black_cell(n, li, lj);
            if (li >= 0) rbk[n] = load_row9(A, F.off(i0 + li - 3, j0 + lj - 3), ld);
        }
#pragma unroll
        for (int m = 0; m < NRS; ++m) {
            const int k = tid + m * NTB;
            const int li = k % (TW + 1) + 2, lj = k / (TW + 1) + 2;
            if (k < (TW + 1) * (TH + 1) && in_dom(F, i0 + li - 3, j0 + lj - 3))
                rrs[m] = load_row9(A, F.off(i0 + li - 3, j0 + lj - 3), ld);
        }
        const int I = i0 / 2 + tid % (TW / 2), J = j0 / 2 + tid / (TW / 2);
        if (tid < (TW / 2) * (TH / 2) && I < C.ni && J < C.nj) {
            T* const* w = P.w;
            const size_t c = C.off(I, J), cw = C.off(I - 1, J), cs = C.off(I, J - 1), csw = C.off(I - 1, J - 1);
            rw[0] = w[W_INJ][c];
            rw[1] = w[W_XW][c];
            rw[2] = w[W_XE][cw];
            rw[3] = w[W_YS][c];
            rw[4] = w[W_YN][cs];
            rw[5] = w[W_SW][c];
            rw[6] = w[W_SE][cw];
            rw[7] = w[W_NW][cs];
            rw[8] = w[W_NE][csw];
        }
    }
    __syncthreads();
    // black on [-2,TW+1) x [-2,TH+1); local parity: black <=> (li + lj) odd
    {
        T nv[NBK];
        int pos[NBK];
#pragma unroll
        for (int n = 0; n < NBK; ++n) {
            int li, lj;
            black_cell(n, li, lj);
            pos[n] = -1;
            nv[n] = T(0);
            if (li < 0) continue;
            const Row9<T> R = PRE ? rbk[n] : load_row9(A, F.off(i0 + li - 3, j0 + lj - 3), ld);
            auto X = [&](int di, int dj) { return sx[lj + dj][li + di]; };
            nv[n] = R.d > T(0) ? (sb[lj][li] + row9_sum(R, nine, X)) / R.d : T(0);
            pos[n] = lj * RW + li;
        }
        __syncthreads();
#pragma unroll
        for (int n = 0; n < NBK; ++n)
            if (pos[n] >= 0) (&sx[0][0])[pos[n]] = nv[n];
    }
    __syncthreads();
    // residual on [-1,TW) x [-1,TH), in place over b (each cell reads only
    // its own b)
#pragma unroll
    for (int m = 0; m < NRS; ++m) {
        const int k = tid + m * NTB;
        if (k >= (TW + 1) * (TH + 1)) continue;
        const int li = k % (TW + 1) + 2, lj = k / (TW + 1) + 2;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        T r = T(0);
        if (in_dom(F, gi, gj)) {
            const Row9<T> R = PRE ? rrs[m] : load_row9(A, F.off(gi, gj), ld);
            auto X = [&](int di, int dj) { return sx[lj + dj][li + di]; };
            r = sb[lj][li] - R.d * sx[lj][li] + row9_sum(R, nine, X);
        }
        sb[lj][li] = r;
    }
    __syncthreads();
    for (int k = tid; k < (TW / 2) * (TH / 2); k += NTB) {
        const int I = i0 / 2 + k % (TW / 2), J = j0 / 2 + k / (TW / 2);
        if (I < C.ni && J < C.nj) {
            const int li = 2 * I - i0 + 3, lj = 2 * J - j0 + 3;
            auto R = [&](int a, int bb) { return sb[lj + bb][li + a]; };
            if constexpr (PRE) {
                // restrict_point with the weights loaded above (same terms, same order)
                T s = rw[0] * R(0, 0);
                s += rw[1] * R(1, 0) + rw[2] * R(-1, 0);
                s += rw[3] * R(0, 1) + rw[4] * R(0, -1);
                s += rw[5] * R(1, 1) + rw[6] * R(-1, 1) + rw[7] * R(1, -1) + rw[8] * R(-1, -1);
                bc[C.off(I, J)] = s;
            } else {
                bc[C.off(I, J)] = restrict_point<T>(C, P, I, J, R);
            }
        }
    }
    for (int k = tid; k < TW * TH; k += NTB) {
        const int li = k % TW + 3, lj = k / TW + 3;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        if (in_dom(F, gi, gj)) xo[F.off(gi, gj)] = sx[lj][li];
    }
}

// up: x = xin + P e, then black and red passes; writes xo
template <typename T, int TW, int TH>
__global__ void __launch_bounds__(NTB) k_up9(Dims F, OpV<T> A, const T* __restrict__ b, const T* __restrict__ xin,
                                             Dims C, Prol<T> P, const T* __restrict__ e, T* __restrict__ xo,
                                             const int* on) {
    if (on && !*on) return;
    constexpr int RW = TW + 4, RH = TH + 4;  // local (li,lj) <-> cell (i0+li-2, j0+lj-2)
    __shared__ T sb[RH][RW];
    __shared__ T sx[RH][RW];
    int tx, ty;
    xcd_tile(tx, ty);
    const int i0 = tx * TW, j0 = ty * TH;
    const int tid = threadIdx.x;
    const size_t ld = F.ld;
    {
        constexpr int BW = RW / 2, BH = RH / 2;
        constexpr int NL2 = (BW * BH + NTB - 1) / NTB;
        const int Ib = i0 / 2 - 1, Jb = j0 / 2 - 1;
#pragma unroll
        for (int m = 0; m < NL2; ++m) {
            const int k = tid + m * NTB;
            if (k >= BW * BH) continue;
            const int a = k % BW, c = k / BW;
            const int I = Ib + a, J = Jb + c;
            T xv[2][2] = {{T(0), T(0)}, {T(0), T(0)}}, bv[2][2] = {{T(0), T(0)}, {T(0), T(0)}};
            if (I >= 0 && J >= 0 && I < C.ni && J < C.nj) {
                const size_t cc = C.off(I, J);
                T* const* w = P.w;
                const T e00 = e[cc], e10 = e[cc + 1], e01 = e[cc + C.ld], e11 = e[cc + C.ld + 1];
                const T pe[2][2] = {
                    {w[W_INJ][cc] * e00, w[W_XW][cc] * e00 + w[W_XE][cc] * e10},
                    {w[W_YS][cc] * e00 + w[W_YN][cc] * e01,
                     w[W_SW][cc] * e00 + w[W_SE][cc] * e10 + w[W_NW][cc] * e01 + w[W_NE][cc] * e11}};
#pragma unroll
                for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                    for (int di = 0; di < 2; ++di) {
                        const int gi = 2 * I + di, gj = 2 * J + dj;
                        if (gi < F.ni && gj < F.nj) {
                            const size_t p = F.off(gi, gj);
                            bv[dj][di] = b[p];
                            xv[dj][di] = xin[p] + pe[dj][di];
                        }
                    }
            }
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    sb[2 * c + dj][2 * a + di] = bv[dj][di];
                    sx[2 * c + dj][2 * a + di] = xv[dj][di];
                }
        }
    }
    // with tiles of at most NTB cells, the couplings of the black cells and
    // of the red tile cell of this thread are loaded before the first barrier
    constexpr bool PRE = TW * TH <= NTB;
    const bool nine = A.cNE != nullptr;
    constexpr int PR = (RW - 2) / 2;
    constexpr int NBK = (PR * (RH - 2) + NTB - 1) / NTB;
    auto black_cell = [&](int m, int& li, int& lj) {
        const int k = tid + m * NTB;
        li = -1;
        lj = 0;
        if (k >= PR * (RH - 2)) return;
        lj = k / PR + 1;
        const int l = 2 * (k % PR) + 1 + (lj & 1);
        if (l >= RW - 1 || !in_dom(F, i0 + l - 2, j0 + lj - 2)) return;
        li = l;
    };
    Row9<T> rbk[PRE ? NBK : 1], rred;
    if constexpr (PRE) {
#pragma unroll
        for (int m = 0; m < NBK; ++m) {
            int li, lj;
            black_cell(m, li, lj);
            if (li >= 0) rbk[m] = load_row9(A, F.off(i0 + li - 2, j0 + lj - 2), ld);
        }
        const int gi = i0 + tid % TW, gj = j0 + tid / TW;
        if (tid < TW * TH && in_dom(F, gi, gj) && ((gi + gj) & 1) == 0) rred = load_row9(A, F.off(gi, gj), ld);
    }
    __syncthreads();
    // black on [-1,TW+1) x [-1,TH+1), staged (Jacobi among black cells)
    {
        T nv[NBK];
        int pos[NBK];
#pragma unroll
        for (int m = 0; m < NBK; ++m) {
            int li, lj;
            black_cell(m, li, lj);
            pos[m] = -1;
            nv[m] = T(0);
            if (li < 0) continue;
            const Row9<T> R = PRE ? rbk[m] : load_row9(A, F.off(i0 + li - 2, j0 + lj - 2), ld);
            auto X = [&](int di, int dj) { return sx[lj + dj][li + di]; };
            nv[m] = R.d > T(0) ? (sb[lj][li] + row9_sum(R, nine, X)) / R.d : T(0);
            pos[m] = lj * RW + li;
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NBK; ++m)
            if (pos[m] >= 0) (&sx[0][0])[pos[m]] = nv[m];
    }
    __syncthreads();
    // red on the tile, written straight out (reads only sx)
    for (int k = tid; k < TW * TH; k += NTB) {
        const int li = k % TW + 2, lj = k / TW + 2;
        const int gi = i0 + li - 2, gj = j0 + lj - 2;
        if (!in_dom(F, gi, gj)) continue;
        const size_t p = F.off(gi, gj);
        T v = sx[lj][li];
        if (((gi + gj) & 1) == 0) {
            const Row9<T> R = PRE ? rred : load_row9(A, p, ld);
            auto X = [&](int di, int dj) { return sx[lj + dj][li + di]; };
            v = R.d > T(0) ? (sb[lj][li] + row9_sum(R, nine, X)) / R.d : T(0);
        }
        xo[p] = v;
    }
}

// ---------------------------------------------------------------------------
// Fine (5-point) level at full size: many tiles hide the memory latency, so
// the operator is read straight from global memory (through the caches)
// rather than staged, which keeps LDS small and occupancy high.
constexpr int GW = 64, GH = 16;  // fine-level tile

// down on the fine level: x = S(0; b) (red from zero, then black), r = b - A
// x, bc = P^T r.  As in the up-sweep, each thread holds the couplings of its
// pairs of tile cells in registers for all three passes; cells of the halo
// read theirs from memory.  SEMI: level 1 coarsens in y only (coarse (I,J)
// on fine (I,2J)).
template <typename T, typename TB, int TW, int TH, bool SEMI>
__global__ void __launch_bounds__(NTB) k_down5(Dims F, OpV<T> A, const TB* __restrict__ b, T* __restrict__ xo,
                                               Dims C, Prol<T> P, T* __restrict__ bc, const int* on) {
    if (on && !*on) return;
    constexpr int RW = TW + 5, RH = TH + 5;  // local (li,lj) <-> cell (i0+li-3, j0+lj-3)
    constexpr int PPR = TW / 2;
    constexpr int PPT = PPR * TH / NTB;
    static_assert(PPR * TH % NTB == 0, "tile pairs must divide among the threads");
    constexpr int NL = (RW * RH + NTB - 1) / NTB;
    __shared__ T sb[RH][RW];
    __shared__ T sx[RH][RW];
    int tx, ty;
    xcd_tile(tx, ty);
    const int i0 = tx * TW, j0 = ty * TH;
    const int tid = threadIdx.x;
    const size_t ld = F.ld;
    T* sxf = &sx[0][0];
    T* sbf = &sb[0][0];
    auto in_tile = [&](int li, int lj) { return li >= 3 && li < TW + 3 && lj >= 3 && lj < TH + 3; };
    auto coef = [&](size_t p, T (&c)[5]) {
        c[0] = A.diag[p];
        c[1] = A.cE[p - 1];
        c[2] = A.cE[p];
        c[3] = A.cN[p - ld];
        c[4] = A.cN[p];
    };
    auto gs = [&](const T (&c)[5], int pos) {
        const T s = c[1] * sxf[pos - 1] + c[2] * sxf[pos + 1] + c[3] * sxf[pos - RW] + c[4] * sxf[pos + RW];
        return c[0] > T(0) ? (sbf[pos] + s) / c[0] : T(0);
    };
    // b; red from a zero guess on the halo (the tile's red cells below)
#pragma unroll
    for (int m = 0; m < NL; ++m) {
        const int k = tid + m * NTB;
        if (k >= RW * RH) continue;
        const int li = k % RW, lj = k / RW;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        T bv = T(0), xv = T(0);
        if (in_dom(F, gi, gj)) {
            const size_t p = F.off(gi, gj);
            bv = (T)b[p];
            if (((gi + gj) & 1) == 0 && !in_tile(li, lj)) {
                const T d = A.diag[p];
                xv = d > T(0) ? bv / d : T(0);
            }
        }
        sb[lj][li] = bv;
        sx[lj][li] = xv;
    }
    // every operand the passes below read from memory is loaded here, before
    // the first barrier: the ring cells' couplings and the restriction
    // weights of the thread's coarse point, beside the pairs' couplings
    constexpr int HW = RW - 2, HH = RH - 2;  // the black ring: li, lj in [1, RW-1) x [1, RH-1) off the tile
    constexpr int NBR = 3 * HW + 3 * TH;     // rows 1, 2, TH+3; columns 1, 2, TW+3 of rows 3..TH+2
    static_assert(NBR <= NTB && TW + 1 + TH <= NTB, "one ring cell per thread");
    int bpos = -1;  // this thread's black ring cell (LDS position), -1 if none
    T bcf[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (tid < NBR) {
        int li, lj;
        if (tid < 3 * HW) {
            li = 1 + tid % HW;
            lj = tid < 2 * HW ? 1 + tid / HW : TH + 3;
        } else {
            const int k2 = tid - 3 * HW;
            li = k2 < 2 * TH ? 1 + k2 % 2 : TW + 3;
            lj = k2 < 2 * TH ? 3 + k2 / 2 : 3 + (k2 - 2 * TH);
        }
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        if (((li + lj) & 1) != 0 && in_dom(F, gi, gj)) {
            coef(F.off(gi, gj), bcf);
            bpos = lj * RW + li;
        }
    }
    int rpos = -1;  // this thread's residual ring cell: row -1 (columns -1..TW-1), column -1 (rows 0..TH-1)
    bool rin = false;
    T rcf[5] = {T(0), T(0), T(0), T(0), T(0)};
    if (tid < TW + 1 + TH) {
        const int li = tid <= TW ? tid + 2 : 2;
        const int lj = tid <= TW ? 2 : tid - TW + 2;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        rpos = lj * RW + li;
        rin = in_dom(F, gi, gj);
        if (rin) coef(F.off(gi, gj), rcf);
    }
    static_assert(SEMI || (TW / 2) * (TH / 2) == NTB, "one coarse point per thread");
    T rw[9];
    int rI = -1, rJ = 0;
    if (!SEMI) {
        const int I = i0 / 2 + tid % (TW / 2), J = j0 / 2 + tid / (TW / 2);
        if (I < C.ni && J < C.nj) {
            T* const* w = P.w;
            const size_t c = C.off(I, J), cw = C.off(I - 1, J), cs = C.off(I, J - 1), csw = C.off(I - 1, J - 1);
            rw[0] = w[W_INJ][c];
            rw[1] = w[W_XW][c];
            rw[2] = w[W_XE][cw];
            rw[3] = w[W_YS][c];
            rw[4] = w[W_YN][cs];
            rw[5] = w[W_SW][c];
            rw[6] = w[W_SE][cw];
            rw[7] = w[W_NW][cs];
            rw[8] = w[W_NE][csw];
            rI = I;
            rJ = J;
        }
    }
    // couplings of the thread's pairs: [0] black, [1] red
    T cf[PPT][2][5];
    int cl[PPT][2];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int k = tid + u * NTB;
        const int lj = k / PPR + 3;
        const int l0 = 2 * (k % PPR) + 3;
        // black <=> (li + lj) odd (i0, j0 even; local offset 3 on both axes)
        const int lb = l0 + (lj & 1), lr = l0 + ((lj + 1) & 1);
        const int ll[2] = {lb, lr};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int gi = i0 + ll[h] - 3, gj = j0 + lj - 3;
            cl[u][h] = -1;
#pragma unroll
            for (int t = 0; t < 5; ++t) cf[u][h][t] = T(0);
            if (!in_dom(F, gi, gj)) continue;
            coef(F.off(gi, gj), cf[u][h]);
            cl[u][h] = lj * RW + ll[h];
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int pos = cl[u][1];
        if (pos >= 0) sxf[pos] = cf[u][1][0] > T(0) ? sbf[pos] / cf[u][1][0] : T(0);
    }
    __syncthreads();
    // black on [-2,TW+1) x [-2,TH+1): the tile's from registers, the halo's
    // from memory (in place: black couples only to red)
#pragma unroll
    for (int u = 0; u < PPT; ++u)
        if (cl[u][0] >= 0) sxf[cl[u][0]] = gs(cf[u][0], cl[u][0]);
    if (bpos >= 0) sxf[bpos] = gs(bcf, bpos);
    __syncthreads();
    // residual on [-1,TW) x [-1,TH), in place over b (each cell reads only
    // its own b)
    auto resid = [&](const T (&c)[5], int pos) {
        return sbf[pos] - c[0] * sxf[pos] +
               (c[1] * sxf[pos - 1] + c[2] * sxf[pos + 1] + c[3] * sxf[pos - RW] + c[4] * sxf[pos + RW]);
    };
    T rt[PPT][2];
#pragma unroll
    for (int u = 0; u < PPT; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h) rt[u][h] = cl[u][h] >= 0 ? resid(cf[u][h], cl[u][h]) : T(0);
    if (rpos >= 0) sbf[rpos] = rin ? resid(rcf, rpos) : T(0);
#pragma unroll
    for (int u = 0; u < PPT; ++u)
#pragma unroll
        for (int h = 0; h < 2; ++h)
            if (cl[u][h] >= 0) sbf[cl[u][h]] = rt[u][h];
    __syncthreads();
    if (SEMI) {
        // P^T r over the fine column: (I,2J) injected, (I,2J+1) by YS(I,J),
        // (I,2J-1) by YN(I,J-1)
        for (int k = tid; k < TW * (TH / 2); k += NTB) {
            const int I = i0 + k % TW, J = j0 / 2 + k / TW;
            if (I < C.ni && J < C.nj) {
                const int li = I - i0 + 3, lj = 2 * J - j0 + 3;
                const size_t cc = C.off(I, J);
                bc[cc] = P.w[W_INJ][cc] * sb[lj][li] + P.w[W_YS][cc] * sb[lj + 1][li] +
                         P.w[W_YN][cc - C.ld] * sb[lj - 1][li];
            }
        }
    } else if (rI >= 0) {
        // restrict_point with the weights loaded above (same terms, same order)
        const int li = 2 * rI - i0 + 3, lj = 2 * rJ - j0 + 3;
        auto R = [&](int a, int bb) { return sb[lj + bb][li + a]; };
        T s = rw[0] * R(0, 0);
        s += rw[1] * R(1, 0) + rw[2] * R(-1, 0);
        s += rw[3] * R(0, 1) + rw[4] * R(0, -1);
        s += rw[5] * R(1, 1) + rw[6] * R(-1, 1) + rw[7] * R(1, -1) + rw[8] * R(-1, -1);
        bc[C.off(rI, rJ)] = s;
    }
    for (int k = tid; k < TW * TH; k += NTB) {
        const int li = k % TW + 3, lj = k / TW + 3;
        const int gi = i0 + li - 3, gj = j0 + lj - 3;
        if (in_dom(F, gi, gj)) xo[F.off(gi, gj)] = sx[lj][li];
    }
}

// up on the fine level: x = xin + P e, black then red, z = x, dots.  The
// tile's cells are taken in pairs (2m, 2m+1) of a row -- one black, one red
// -- and each thread loads the couplings of its pairs once into registers:
// the black pass uses the black cell's, the red pass the red cell's, so no
// coefficient is read from memory twice.  The one-cell black ring around the
// tile is smoothed from global reads.  b stays in LDS in FP64 for the dots.
// SEMI: level 1 coarsens in y only.
template <typename T, typename V, int TW, int TH, bool SEMI>
__global__ void __launch_bounds__(NTB) k_up5(Dims F, OpV<T> A, const V* __restrict__ b,
                                             const T* __restrict__ xin, Dims C, Prol<T> P, const T* __restrict__ e,
                                             V* __restrict__ z, const V* __restrict__ q, Slots Sl,
                                             const int* on) {
    if (on && !*on) return;
    constexpr int RW = TW + 4, RH = TH + 4;  // local (li,lj) <-> cell (i0+li-2, j0+lj-2)
    constexpr int PPR = TW / 2;              // pairs per row
    constexpr int PPT = PPR * TH / NTB;      // pairs per thread
    static_assert(PPR * TH % NTB == 0, "tile pairs must divide among the threads");
    __shared__ double sb[RH][RW];
    __shared__ T sx[RH][RW];
    int tx, ty;
    xcd_tile(tx, ty);
    const int i0 = tx * TW, j0 = ty * TH;
    const int tid = threadIdx.x;
    const size_t ld = F.ld;
    // b and x + P e per 2x2 block of cells sharing one coarse parent (the
    // region starts on an even cell, so blocks align with coarse cells); with
    // SEMI per 1x2 block (fine (I,2J) and (I,2J+1) of coarse (I,J))
    if (SEMI) {
        constexpr int BW = RW, BH = RH / 2;
        constexpr int NL2 = (BW * BH + NTB - 1) / NTB;
        const int Ib = i0 - 2, Jb = j0 / 2 - 1;
#pragma unroll
        for (int m = 0; m < NL2; ++m) {
            const int k = tid + m * NTB;
            if (k >= BW * BH) continue;
            const int a = k % BW, cr = k / BW;
            const int I = Ib + a, J = Jb + cr;
            T xv[2] = {T(0), T(0)};
            double bv[2] = {0.0, 0.0};
            if (I >= 0 && J >= 0 && I < C.ni && J < C.nj) {
                const size_t cc = C.off(I, J);
                T* const* w = P.w;
                const T e00 = e[cc], e01 = e[cc + C.ld];
                const T pe[2] = {w[W_INJ][cc] * e00, w[W_YS][cc] * e00 + w[W_YN][cc] * e01};
#pragma unroll
                for (int dj = 0; dj < 2; ++dj) {
                    const int gj = 2 * J + dj;
                    if (gj < F.nj) {
                        const size_t p = F.off(I, gj);
                        bv[dj] = (double)b[p];
                        xv[dj] = xin[p] + pe[dj];
                    }
                }
            }
#pragma unroll
            for (int dj = 0; dj < 2; ++dj) {
                sb[2 * cr + dj][a] = bv[dj];
                sx[2 * cr + dj][a] = xv[dj];
            }
        }
    } else {
        constexpr int BW = RW / 2, BH = RH / 2;
        constexpr int NL2 = (BW * BH + NTB - 1) / NTB;
        const int Ib = i0 / 2 - 1, Jb = j0 / 2 - 1;
#pragma unroll
        for (int m = 0; m < NL2; ++m) {
            const int k = tid + m * NTB;
            if (k >= BW * BH) continue;
            const int a = k % BW, c = k / BW;
            const int I = Ib + a, J = Jb + c;
            T xv[2][2] = {{T(0), T(0)}, {T(0), T(0)}};
            double bv[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
            if (I >= 0 && J >= 0 && I < C.ni && J < C.nj) {
                const size_t cc = C.off(I, J);
                T* const* w = P.w;
                const T e00 = e[cc], e10 = e[cc + 1], e01 = e[cc + C.ld], e11 = e[cc + C.ld + 1];
                const T pe[2][2] = {
                    {w[W_INJ][cc] * e00, w[W_XW][cc] * e00 + w[W_XE][cc] * e10},
                    {w[W_YS][cc] * e00 + w[W_YN][cc] * e01,
                     w[W_SW][cc] * e00 + w[W_SE][cc] * e10 + w[W_NW][cc] * e01 + w[W_NE][cc] * e11}};
#pragma unroll
                for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                    for (int di = 0; di < 2; ++di) {
                        const int gi = 2 * I + di, gj = 2 * J + dj;
                        if (gi < F.ni && gj < F.nj) {
                            const size_t p = F.off(gi, gj);
                            bv[dj][di] = (double)b[p];
                            xv[dj][di] = xin[p] + pe[dj][di];
                        }
                    }
            }
#pragma unroll
            for (int dj = 0; dj < 2; ++dj)
#pragma unroll
                for (int di = 0; di < 2; ++di) {
                    sb[2 * c + dj][2 * a + di] = bv[dj][di];
                    sx[2 * c + dj][2 * a + di] = xv[dj][di];
                }
        }
    }
    // couplings of the thread's pairs: [0] the black cell, [1] the red one
    // (d, west, east, south, north)
    T cf[PPT][2][5];
    int cl[PPT][2];  // LDS position (lj * RW + li), -1 outside the domain
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
        const int k = tid + u * NTB;
        const int lj = k / PPR + 2;
        const int l0 = 2 * (k % PPR) + 2;
        const int lb = l0 + ((lj + 1) & 1), lr = l0 + (lj & 1);  // black <=> (li + lj) odd
        const int ll[2] = {lb, lr};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int gi = i0 + ll[h] - 2, gj = j0 + lj - 2;
            cl[u][h] = -1;
#pragma unroll
            for (int t = 0; t < 5; ++t) cf[u][h][t] = T(0);
            if (!in_dom(F, gi, gj)) continue;
            const size_t p = F.off(gi, gj);
            cf[u][h][0] = A.diag[p];
            cf[u][h][1] = A.cE[p - 1];
            cf[u][h][2] = A.cE[p];
            cf[u][h][3] = A.cN[p - ld];
            cf[u][h][4] = A.cN[p];
            cl[u][h] = lj * RW + ll[h];
        }
    }
    __syncthreads();
    T* sxf = &sx[0][0];
    const double* sbf = &sb[0][0];
    auto gs = [&](const T (&c)[5], int pos) {
        const T s = c[1] * sxf[pos - 1] + c[2] * sxf[pos + 1] + c[3] * sxf[pos - RW] + c[4] * sxf[pos + RW];
        return c[0] > T(0) ? ((T)sbf[pos] + s) / c[0] : T(0);
    };
    // black: the tile's black cells from registers, the ring from memory (in
    // place: black couples only to red)
#pragma unroll
    for (int u = 0; u < PPT; ++u)
        if (cl[u][0] >= 0) sxf[cl[u][0]] = gs(cf[u][0], cl[u][0]);
    {
        // ring cells: rows -1 and TH (columns -1..TW), columns -1 and TW (rows 0..TH-1)
        constexpr int NRING = 2 * (TW + 2) + 2 * TH;
        for (int k = tid; k < NRING; k += NTB) {
            int li, lj;
            if (k < TW + 2) {
                li = k + 1;
                lj = 1;
            } else if (k < 2 * (TW + 2)) {
                li = k - (TW + 2) + 1;
                lj = TH + 2;
            } else {
                const int r = k - 2 * (TW + 2);
                li = (r & 1) ? TW + 2 : 1;
                lj = r / 2 + 2;
            }
            if (((li + lj) & 1) == 0) continue;  // red
            const int gi = i0 + li - 2, gj = j0 + lj - 2;
            if (!in_dom(F, gi, gj)) continue;
            const size_t p = F.off(gi, gj);
            const T c[5] = {A.diag[p], A.cE[p - 1], A.cE[p], A.cN[p - ld], A.cN[p]};
            sxf[lj * RW + li] = gs(c, lj * RW + li);
        }
    }
    __syncthreads();
    // red from registers; z = x, dots (b, z), (z, q)
    double v0 = 0.0, v1 = 0.0;
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int pos = cl[u][h];
            if (pos < 0) continue;
            const T v = h ? gs(cf[u][1], pos) : sxf[pos];
            const int li = pos % RW, lj = pos / RW;
            const size_t p = F.off(i0 + li - 2, j0 + lj - 2);
            const V zs = (V)v;
            const double zv = (double)zs;
            z[p] = zs;
            v0 += sbf[pos] * zv;
            v1 += zv * (double)q[p];
        }
    }
    double vv[2] = {v0, v1};
    block_partials<2, NTB>(vv, Sl.part, Sl.max_blocks);
}

// the fine-level kernels with the level-1 coarsening chosen at run time
template <typename T, typename TB>
void launch_down5(bool semi, hipStream_t s, Dims F, OpV<T> A, const TB* b, T* xo, Dims C, Prol<T> P, T* bc,
                  const int* on) {
    const dim3 g((F.ni + GW - 1) / GW, (F.nj + GH - 1) / GH);
    if (semi)
        hipLaunchKernelGGL((k_down5<T, TB, GW, GH, true>), g, dim3(NTB), 0, s, F, A, b, xo, C, P, bc, on);
    else
        hipLaunchKernelGGL((k_down5<T, TB, GW, GH, false>), g, dim3(NTB), 0, s, F, A, b, xo, C, P, bc, on);
}

template <typename T, typename V>
void launch_up5(bool semi, hipStream_t s, Dims F, OpV<T> A, const V* b, const T* xin, Dims C, Prol<T> P,
                const T* e, V* z, const V* q, Slots S, const int* on) {
    const dim3 g((F.ni + GW - 1) / GW, (F.nj + GH - 1) / GH);
    if (semi)
        hipLaunchKernelGGL((k_up5<T, V, GW, GH, true>), g, dim3(NTB), 0, s, F, A, b, xin, C, P, e, z, q, S, on);
    else
        hipLaunchKernelGGL((k_up5<T, V, GW, GH, false>), g, dim3(NTB), 0, s, F, A, b, xin, C, P, e, z, q, S, on);
}

// ---------------------------------------------------------------------------
// tail: the V-cycle on levels lt..nlev-1 in one workgroup, all of it in LDS
// (compact layout cd: ld = ni+1; the arrays of a level follow each other)
enum { TA_D = 0, TA_E, TA_N, TA_NE, TA_NW, TA_X, TA_B, TA_R, TA_COUNT };

template <typename T>
struct TailLevel {
    Dims d;   // global layout
    Dims cd;  // LDS layout
    int base;  // LDS offset (elements) of this level's TA_COUNT arrays
    int wbase; // LDS offset of its 9 prolongation-weight arrays (next level's cd)
    int sh;    // thread k of a pass takes cell (k & (2^sh - 1), k >> sh), 2^sh >= ni
    OpV<T> A;
    Prol<T> P;  // to the next level
    T* x;       // global x of the level (written back for the first tail level)
    const T* b; // global b (read for the first tail level)
};

template <typename T>
struct TailArgs {
    int n;  // number of tail levels (the last is the coarsest)
    TailLevel<T> L[MAX_TAIL];
    const double* inv;  // dense inverse of the coarsest operator (FP64 in both precisions)
    const int* on; // the PCG's convergence flag (the cycle runs while *on), or null
    int krylov;    // inner flexible-CG iterations on the first tail level (0: one V-cycle)
    int pbase;     // LDS offset of the inner iteration's direction vector
    int invbase;   // LDS offset (in T) of the staged FP64 dense inverse, 8-byte aligned
};

// offset of cell (i,j) in a level's LDS arrays (compact layout, 32-bit)
__device__ inline int toff(const Dims& D, int i, int j) { return (j + 2) * D.ld + i; }

template <typename T>
__device__ inline T* tarr(T* lds, const TailLevel<T>& L, int k) {
    return lds + L.base + (size_t)k * L.cd.alloc_elems();
}

// One colour pass of the smoother, double-buffered: the cells of colour
// `col` get (b + sum of couplings * src) / d, the others keep src; every
// cell is written to dst, so one barrier ends the pass (same-colour diagonal
// couplings read src, i.e. Jacobi-wise).  zero_start: src = 0.
template <typename T>
__device__ void tail_pass(T* lds, const TailLevel<T>& L, const T* src, T* dst, int col, bool zero_start) {
    const Dims& D = L.cd;
    const int ld = D.ld;
    const T* d = tarr(lds, L, TA_D);
    const T* e = tarr(lds, L, TA_E);
    const T* nn = tarr(lds, L, TA_N);
    const T* ne = tarr(lds, L, TA_NE);
    const T* nw = tarr(lds, L, TA_NW);
    const T* b = tarr(lds, L, TA_B);
    const int mask = (1 << L.sh) - 1;
    for (int k = threadIdx.x; k < (D.nj << L.sh); k += TAIL_THREADS) {
        const int i = k & mask, j = k >> L.sh;
        if (i >= D.ni) continue;
        const int p = toff(D, i, j);
        T v = zero_start ? T(0) : src[p];
        if (((i + j) & 1) == col) {
            const T dd = d[p];
            if (zero_start) {
                v = dd > T(0) ? b[p] / dd : T(0);
            } else {
                const T s = e[p - 1] * src[p - 1] + e[p] * src[p + 1] + nn[p - ld] * src[p - ld] +
                            nn[p] * src[p + ld] + ne[p] * src[p + ld + 1] + ne[p - ld - 1] * src[p - ld - 1] +
                            nw[p] * src[p + ld - 1] + nw[p - ld + 1] * src[p - ld + 1];
                v = dd > T(0) ? (b[p] + s) / dd : T(0);
            }
        }
        dst[p] = v;
    }
    __syncthreads();
}

// The first tail level's couplings of the cell a thread owns in every pass,
// held in registers for the whole kernel (used when each thread has at most
// one cell of that level: (nj << sh) <= TAIL_THREADS)
template <typename T>
struct OwnRow {
    bool own;  // the thread has a cell
    int col;   // its colour, (i + j) & 1
    int p;     // its LDS offset
    T d, ew, ee, ns, nn, ne, nesw, nw, nwse;  // d[p], e[p-1], e[p], n[p-ld], n[p], ne[p], ne[p-ld-1], nw[p], nw[p-ld+1]
};

template <typename T>
__device__ OwnRow<T> own_row(T* lds, const TailLevel<T>& L) {
    OwnRow<T> R{};
    const Dims& D = L.cd;
    const int k = threadIdx.x, i = k & ((1 << L.sh) - 1), j = k >> L.sh;
    R.own = k < (D.nj << L.sh) && i < D.ni;
    if (!R.own) return R;
    const int ld = D.ld, p = toff(D, i, j);
    const T* d = tarr(lds, L, TA_D);
    const T* e = tarr(lds, L, TA_E);
    const T* nn = tarr(lds, L, TA_N);
    const T* ne = tarr(lds, L, TA_NE);
    const T* nw = tarr(lds, L, TA_NW);
    R.col = (i + j) & 1;
    R.p = p;
    R.d = d[p];
    R.ew = e[p - 1];
    R.ee = e[p];
    R.ns = nn[p - ld];
    R.nn = nn[p];
    R.ne = ne[p];
    R.nesw = ne[p - ld - 1];
    R.nw = nw[p];
    R.nwse = nw[p - ld + 1];
    return R;
}

// tail_pass on the first level with the couplings from registers (same
// terms in the same order)
template <typename T>
__device__ void tail_pass_own(T* lds, const TailLevel<T>& L, const OwnRow<T>& R, const T* src, T* dst, int col,
                              bool zero_start) {
    if (R.own) {
        const int ld = L.cd.ld, p = R.p;
        const T* b = tarr(lds, L, TA_B);
        T v = zero_start ? T(0) : src[p];
        if (R.col == col) {
            const T dd = R.d;
            if (zero_start) {
                v = dd > T(0) ? b[p] / dd : T(0);
            } else {
                const T s = R.ew * src[p - 1] + R.ee * src[p + 1] + R.ns * src[p - ld] + R.nn * src[p + ld] +
                            R.ne * src[p + ld + 1] + R.nesw * src[p - ld - 1] + R.nw * src[p + ld - 1] +
                            R.nwse * src[p - ld + 1];
                v = dd > T(0) ? (b[p] + s) / dd : T(0);
            }
        }
        dst[p] = v;
    }
    __syncthreads();
}

// stage operators, weights and the first level's b into LDS
template <typename T>
__device__ void tail_stage(T* lds, const TailArgs<T>& a) {
    const int tid = threadIdx.x;
    // stage operators, weights and the first level's b
    for (int l = 0; l < a.n; ++l) {
        const TailLevel<T>& L = a.L[l];
        const int n = L.d.ni * L.d.nj;
        const bool nine = L.A.cNE != nullptr;
        for (int k = tid; k < n; k += TAIL_THREADS) {
            const int i = k % L.d.ni, j = k / L.d.ni;
            const size_t g = L.d.off(i, j), c = L.cd.off(i, j);
            tarr(lds, L, TA_D)[c] = L.A.diag[g];
            tarr(lds, L, TA_E)[c] = L.A.cE[g];
            tarr(lds, L, TA_N)[c] = L.A.cN[g];
            tarr(lds, L, TA_NE)[c] = nine ? L.A.cNE[g] : T(0);
            tarr(lds, L, TA_NW)[c] = nine ? L.A.cNW[g] : T(0);
            if (l == 0) tarr(lds, L, TA_B)[c] = L.b[g];
        }
        if (l + 1 < a.n) {
            const TailLevel<T>& Cl = a.L[l + 1];
            const int nc = Cl.d.ni * Cl.d.nj;
            for (int k = tid; k < nc; k += TAIL_THREADS) {
                const int I = k % Cl.d.ni, J = k / Cl.d.ni;
                const size_t g = Cl.d.off(I, J), c = Cl.cd.off(I, J);
#pragma unroll
                for (int m = 0; m < 9; ++m) lds[L.wbase + (size_t)m * Cl.cd.alloc_elems() + c] = L.P.w[m][g];
            }
        }
    }
}

// One V-cycle over the tail levels: level 0's TA_B -> level 0's TA_R.  On
// every level the smoothed x ends in TA_R; TA_X holds the first pass and then
// the residual that is restricted.
// With `regs`, level 0 takes its couplings from R (each thread's cell).
template <typename T>
__device__ void tail_vcycle(T* lds, const TailArgs<T>& a, const OwnRow<T>& R, bool regs) {
    const int tid = threadIdx.x;
    const int last = a.n - 1;
    for (int l = 0; l < last; ++l) {
        const TailLevel<T>& L = a.L[l];
        const TailLevel<T>& Cl = a.L[l + 1];
        const Dims& D = L.cd;
        const int ld = D.ld;
        T* xa = tarr(lds, L, TA_X);
        T* xb = tarr(lds, L, TA_R);
        const T* b = tarr(lds, L, TA_B);
        const T* x = xb;
        T* r = xa;
        if (l == 0 && regs) {
            tail_pass_own<T>(lds, L, R, nullptr, xa, 0, true);
            tail_pass_own<T>(lds, L, R, xa, xb, 1, false);
            if (R.own) {
                const int p = R.p;
                const T s = R.ew * x[p - 1] + R.ee * x[p + 1] + R.ns * x[p - ld] + R.nn * x[p + ld] +
                            R.ne * x[p + ld + 1] + R.nesw * x[p - ld - 1] + R.nw * x[p + ld - 1] +
                            R.nwse * x[p - ld + 1];
                r[p] = b[p] - R.d * x[p] + s;
            }
        } else {
            tail_pass<T>(lds, L, nullptr, xa, 0, true);
            tail_pass<T>(lds, L, xa, xb, 1, false);
            const T* d = tarr(lds, L, TA_D);
            const T* e = tarr(lds, L, TA_E);
            const T* nn = tarr(lds, L, TA_N);
            const T* ne = tarr(lds, L, TA_NE);
            const T* nw = tarr(lds, L, TA_NW);
            const int mask = (1 << L.sh) - 1;
            for (int k = tid; k < (D.nj << L.sh); k += TAIL_THREADS) {
                const int i = k & mask;
                if (i >= D.ni) continue;
                const int p = toff(D, i, k >> L.sh);
                const T s = e[p - 1] * x[p - 1] + e[p] * x[p + 1] + nn[p - ld] * x[p - ld] + nn[p] * x[p + ld] +
                            ne[p] * x[p + ld + 1] + ne[p - ld - 1] * x[p - ld - 1] + nw[p] * x[p + ld - 1] +
                            nw[p - ld + 1] * x[p - ld + 1];
                r[p] = b[p] - d[p] * x[p] + s;
            }
        }
        __syncthreads();
        T* bc = tarr(lds, Cl, TA_B);
        const Dims& Cd = Cl.cd;
        const int ws = (int)Cd.alloc_elems(), cld = Cd.ld;
        const T* w0 = lds + L.wbase;
        const int cmask = (1 << Cl.sh) - 1;
        for (int k = tid; k < (Cd.nj << Cl.sh); k += TAIL_THREADS) {
            const int I = k & cmask, J = k >> Cl.sh;
            if (I >= Cd.ni) continue;
            const int c = toff(Cd, I, J), f = toff(D, 2 * I, 2 * J);
            const T* w = w0 + c;
            T s = w[W_INJ * ws] * r[f];
            s += w[W_XW * ws] * r[f + 1] + w[W_XE * ws - 1] * r[f - 1];
            s += w[W_YS * ws] * r[f + ld] + w[W_YN * ws - cld] * r[f - ld];
            s += w[W_SW * ws] * r[f + ld + 1] + w[W_SE * ws - 1] * r[f + ld - 1] + w[W_NW * ws - cld] * r[f - ld + 1] +
                 w[W_NE * ws - cld - 1] * r[f - ld - 1];
            bc[c] = s;
        }
        __syncthreads();
    }
    {
        // dense solve on the coarsest level in FP64 (in both precisions):
        // four threads per row, the inverse staged in LDS
        const TailLevel<T>& L = a.L[last];
        const Dims& D = L.cd;
        const int n = D.ni * D.nj;
        const T* b = tarr(lds, L, TA_B);
        const double* inv = reinterpret_cast<const double*>(lds + a.invbase);
        T* x = tarr(lds, L, TA_R);
        const int row = tid >> 2, quarter = tid & 3;
        double s = 0.0;
        if (row < n)
            for (int q = quarter; q < n; q += 4) s += inv[row * n + q] * (double)b[D.off(q % D.ni, q / D.ni)];
        s += __shfl_xor(s, 1, 64);
        s += __shfl_xor(s, 2, 64);
        if (row < n && quarter == 0) x[D.off(row % D.ni, row / D.ni)] = (T)s;
        __syncthreads();
    }
    for (int l = last - 1; l >= 0; --l) {
        const TailLevel<T>& L = a.L[l];
        const TailLevel<T>& Cl = a.L[l + 1];
        const Dims& D = L.cd;
        const Dims& Cd = Cl.cd;
        const int ws = (int)Cd.alloc_elems(), cld = Cd.ld;
        const T* w0 = lds + L.wbase;
        const T* ec = tarr(lds, Cl, TA_R);
        T* xa = tarr(lds, L, TA_X);
        T* xb = tarr(lds, L, TA_R);
        const int mask = (1 << L.sh) - 1;
        for (int k = tid; k < (D.nj << L.sh); k += TAIL_THREADS) {
            const int i = k & mask, j = k >> L.sh;
            if (i >= D.ni) continue;
            const int c = toff(Cd, i >> 1, j >> 1);
            const T* w = w0 + c;
            const int pi = i & 1, pj = j & 1;
            T v;
            if (!pi && !pj) v = w[W_INJ * ws] * ec[c];
            else if (pi && !pj) v = w[W_XW * ws] * ec[c] + w[W_XE * ws] * ec[c + 1];
            else if (!pi && pj) v = w[W_YS * ws] * ec[c] + w[W_YN * ws] * ec[c + cld];
            else v = w[W_SW * ws] * ec[c] + w[W_SE * ws] * ec[c + 1] + w[W_NW * ws] * ec[c + cld] +
                     w[W_NE * ws] * ec[c + cld + 1];
            xb[toff(D, i, j)] += v;
        }
        __syncthreads();
        if (l == 0 && regs) {
            tail_pass_own<T>(lds, L, R, xb, xa, 1, false);
            tail_pass_own<T>(lds, L, R, xa, xb, 0, false);
        } else {
            tail_pass<T>(lds, L, xb, xa, 1, false);
            tail_pass<T>(lds, L, xa, xb, 0, false);
        }
    }
}

// Sum of v over the workgroup, returned to every thread (part: 16 slots of
// LDS per value, not reused before a barrier)
template <int NV>
__device__ inline void tail_allreduce(double (&v)[NV], double (*part)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) part[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int w = 0; w < TAIL_THREADS / 64; ++w) s += part[w][k];
        v[k] = s;
    }
}

// The tail: one V-cycle over levels lt..nlev-1, or -- with a.krylov > 0 --
// that many iterations of flexible CG on level lt preconditioned by the
// V-cycle (x0 = 0; each direction A-orthogonal to the previous one).  The
// coarse levels of a bubbly flow carry slow modes (the pressure of each
// bubble against the liquid around it) that one V-cycle resolves poorly;
// the inner iterations remove them at the cost of a few V-cycles of the
// smallest levels, all inside one workgroup.
template <typename T>
__global__ void __launch_bounds__(TAIL_THREADS) k_tail(TailArgs<T> a, int lds_elems) {
    if (a.on && !*a.on) return;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T* lds = reinterpret_cast<T*>(smem_raw);
    const int tid = threadIdx.x;
    for (int k = tid; k < lds_elems; k += TAIL_THREADS) lds[k] = T(0);
    __syncthreads();
    tail_stage<T>(lds, a);
    {
        const Dims& Dc = a.L[a.n - 1].d;
        const int nc = Dc.ni * Dc.nj;
        double* inv = reinterpret_cast<double*>(lds + a.invbase);
        for (int k = tid; k < nc * nc; k += TAIL_THREADS) inv[k] = a.inv[k];
    }
    __syncthreads();
    const TailLevel<T>& L = a.L[0];
    const Dims& D = L.cd;
    const int n = D.ni * D.nj;
    // level 0's couplings in registers when each thread has at most one cell
    const bool regs = (D.nj << L.sh) <= TAIL_THREADS;
    const OwnRow<T> R = regs ? own_row<T>(lds, L) : OwnRow<T>{};
    if (a.krylov <= 0) {
        tail_vcycle<T>(lds, a, R, regs);
        const T* x = tarr(lds, L, TA_R);
        for (int k = tid; k < n; k += TAIL_THREADS) {
            const int i = k % D.ni, j = k / D.ni;
            L.x[L.d.off(i, j)] = x[D.off(i, j)];
        }
        return;
    }
    // inner flexible CG; level lt has at most TAIL_THREADS points, one per thread
    __shared__ double part_b[TAIL_THREADS / 64][2];
    __shared__ double part_a[TAIL_THREADS / 64][2];
    const size_t ld = D.ld;
    const bool own = tid < n;
    const size_t p = own ? D.off(tid % D.ni, tid / D.ni) : 0;
    T* res = tarr(lds, L, TA_B);  // the V-cycle's right-hand side
    const T* z = tarr(lds, L, TA_R);
    T* pv = lds + a.pbase;
    const T* d = tarr(lds, L, TA_D);
    const T* e = tarr(lds, L, TA_E);
    const T* nn = tarr(lds, L, TA_N);
    const T* ne = tarr(lds, L, TA_NE);
    const T* nw = tarr(lds, L, TA_NW);
    T xacc = T(0), qv = T(0);
    double pq_prev = 0.0, rr0 = 0.0;
    for (int it = 0; it < a.krylov; ++it) {
        tail_vcycle<T>(lds, a, R, regs);  // ends with a barrier
        const T zp = own ? z[p] : T(0);
        const double rp = own ? (double)res[p] : 0.0;
        double v[2] = {(double)zp * (double)qv, rp * rp};
        tail_allreduce<2>(v, part_b);
        // every thread holds the same sums, so the exits below are uniform:
        // stop once the inner residual has fallen 1e4 (further iterations
        // would work on rounding noise, in FP32 above all)
        if (it == 0) rr0 = v[1];
        else if (!(v[1] > 1e-8 * rr0)) break;
        const T beta = (it > 0 && pq_prev > 0.0) ? (T)(-v[0] / pq_prev) : T(0);
        if (own) pv[p] = zp + beta * pv[p];
        __syncthreads();
        T pp = T(0);
        if (own) {
            pp = pv[p];
            const T s = e[p - 1] * pv[p - 1] + e[p] * pv[p + 1] + nn[p - ld] * pv[p - ld] + nn[p] * pv[p + ld] +
                        ne[p] * pv[p + ld + 1] + ne[p - ld - 1] * pv[p - ld - 1] + nw[p] * pv[p + ld - 1] +
                        nw[p - ld + 1] * pv[p - ld + 1];
            qv = d[p] * pp - s;
        }
        double v2[2] = {(double)pp * rp, (double)pp * (double)qv};
        tail_allreduce<2>(v2, part_a);
        const double al = v2[0] / v2[1];
        if (!(v2[1] > 0.0) || !(fabs(al) < 1e30)) break;
        const T alpha = (T)al;
        pq_prev = v2[1];
        if (own) {
            xacc += alpha * pp;
            res[p] -= alpha * qv;
        }
        __syncthreads();
    }
    if (own) L.x[L.d.off(tid % D.ni, tid / D.ni)] = xacc;
}

// coarsest-only hierarchy (grids of at most COARSEST_MAX cells): the dense
// FP64 solve
template <typename TB, typename TZ>
__global__ void __launch_bounds__(256) k_direct(Dims C, const double* __restrict__ inv, const TB* __restrict__ b,
                                                TZ* __restrict__ x, const int* on) {
    if (on && !*on) return;
    const int n = C.ni * C.nj;
    const int r = threadIdx.x;
    if (r >= n) return;
    double s = 0.0;
    for (int q = 0; q < n; ++q) s += inv[(size_t)r * n + q] * (double)b[C.off(q % C.ni, q / C.ni)];
    x[C.off(r % C.ni, r / C.ni)] = (TZ)s;
}

template <typename T>
struct Acc;
template <>
struct Acc<double> {
    static const Op<double>& op(const Level& L) { return L.A; }
    static const Prol<double>& pr(const Level& L) { return L.P; }
    static const LevelVecs<double>& vec(const Level& L) { return L.v; }
};
template <>
struct Acc<float> {
    static const Op<float>& op(const Level& L) { return L.Af; }
    static const Prol<float>& pr(const Level& L) { return L.Pf; }
    static const LevelVecs<float>& vec(const Level& L) { return L.vf; }
};

inline dim3 tiles(const Dims& d, int tw, int th) { return dim3((d.ni + tw - 1) / tw, (d.nj + th - 1) / th); }

// 64x16 tiles on coarse levels of 2048^2 and more, 32x8 below: the latency
// of a tile's passes is hidden only with several tiles per CU
inline bool big(const Dims& d) { return (size_t)d.ni * d.nj >= (size_t)2048 * 2048; }

inline Dims compact(const Dims& d) {
    Dims c = d;
    c.ld = d.ni + 1;
    return c;
}

// z = M^{-1} r ; fine-level dots (r,z), (z,q) go to the partials slot S;
// V = type of the PCG vectors r, z, q; every kernel returns at once while
// *on == 0 (a converged PCG)
template <typename T, typename V>
int vcycle(Ctx* c, const V* r, V* z, const V* q, Slots S, int* dots_done, const int* on) {
    hipStream_t s = c->stream;
    const int nl = c->nlev;
    *dots_done = 0;
    if (nl == 1) {
        const Level& L = c->lev[0];
        hipLaunchKernelGGL((k_direct<V, V>), dim3(1), dim3(256), 0, s, L.d, (const double*)c->coarse_inv, r, z, on);
        SG_HIP(hipGetLastError());
        return SG_OK;
    }
    *dots_done = 1;
    int lt = 1;  // first level of the tail
    while (lt < nl - 1 && (size_t)c->lev[lt].d.ni * c->lev[lt].d.nj > (size_t)TAIL_MAX) ++lt;
    // down
    for (int l = 0; l < lt; ++l) {
        const Level& L = c->lev[l];
        const Level& Cl = c->lev[l + 1];
        if (l == 0)
            launch_down5<T, V>(c->semi0 != 0, s, L.d, opv(Acc<T>::op(L)), r, Acc<T>::vec(L).x, Cl.d, Acc<T>::pr(L),
                               Acc<T>::vec(Cl).b, on);
        else if (big(L.d))
            hipLaunchKernelGGL((k_down9<T, 64, 16>), tiles(L.d, 64, 16), dim3(NTB), 0, s, L.d, opv(Acc<T>::op(L)),
                               (const T*)Acc<T>::vec(L).b, Acc<T>::vec(L).x, Cl.d, Acc<T>::pr(L),
                               Acc<T>::vec(Cl).b, on);
        else
            hipLaunchKernelGGL((k_down9<T, 32, 8>), tiles(L.d, 32, 8), dim3(NTB), 0, s, L.d, opv(Acc<T>::op(L)),
                               (const T*)Acc<T>::vec(L).b, Acc<T>::vec(L).x, Cl.d, Acc<T>::pr(L),
                               Acc<T>::vec(Cl).b, on);
    }
    // tail
    TailArgs<T> ta{};
    ta.n = nl - lt;
    if (ta.n > MAX_TAIL) {
        set_error("multigrid: too many tail levels");
        return SG_ERR_STATE;
    }
    int off = 0;
    for (int k = 0; k < ta.n; ++k) {
        const Level& L = c->lev[lt + k];
        TailLevel<T>& t = ta.L[k];
        t.d = L.d;
        t.cd = compact(L.d);
        t.sh = 0;
        while ((1 << t.sh) < L.d.ni) ++t.sh;
        t.base = off;
        off += TA_COUNT * (int)t.cd.alloc_elems();
        t.A = opv(Acc<T>::op(L));
        t.P = Acc<T>::pr(L);
        t.x = Acc<T>::vec(L).x;
        t.b = Acc<T>::vec(L).b;
        if (k + 1 < ta.n) {
            t.wbase = off;
            off += 9 * (int)compact(c->lev[lt + k + 1].d).alloc_elems();
        }
    }
    ta.krylov = c->prm.mg_coarse_krylov;
    ta.pbase = off;
    if (ta.krylov > 0) {
        if (c->lev[lt].d.ni * c->lev[lt].d.nj > TAIL_THREADS) {
            set_error("multigrid: the first tail level exceeds one point per thread");
            return SG_ERR_STATE;
        }
        off += (int)compact(c->lev[lt].d).alloc_elems();
    }
    // the FP64 inverse: 8-byte aligned, sizeof(double) / sizeof(T) slots per entry
    constexpr int per = (int)(sizeof(double) / sizeof(T));
    off = (off + per - 1) / per * per;
    ta.invbase = off;
    off += per * (int)(c->lev[nl - 1].d.ni * c->lev[nl - 1].d.nj) * (int)(c->lev[nl - 1].d.ni * c->lev[nl - 1].d.nj);
    const size_t lds_bytes = (size_t)off * sizeof(T);
    if (lds_bytes > TAIL_LDS_MAX) {
        set_error("multigrid: tail levels exceed the LDS budget");
        return SG_ERR_STATE;
    }
    ta.inv = c->coarse_inv;
    ta.on = on;
    hipLaunchKernelGGL((k_tail<T>), dim3(1), dim3(TAIL_THREADS), lds_bytes, s, ta, off);
    // up: level l reads x_l from the down pass and e_{l+1}; writes its own
    // correction into r_l (its solution for the level above)
    for (int l = lt - 1; l >= 0; --l) {
        const Level& L = c->lev[l];
        const Level& Cl = c->lev[l + 1];
        const T* e = (l + 1 == lt) ? Acc<T>::vec(Cl).x : Acc<T>::vec(Cl).r;
        if (l == 0)
            launch_up5<T, V>(c->semi0 != 0, s, L.d, opv(Acc<T>::op(L)), r, (const T*)Acc<T>::vec(L).x, Cl.d,
                             Acc<T>::pr(L), e, z, q, S, on);
        else if (big(L.d))
            hipLaunchKernelGGL((k_up9<T, 64, 16>), tiles(L.d, 64, 16), dim3(NTB), 0, s, L.d, opv(Acc<T>::op(L)),
                               (const T*)Acc<T>::vec(L).b, (const T*)Acc<T>::vec(L).x, Cl.d, Acc<T>::pr(L), e,
                               Acc<T>::vec(L).r, on);
        else
            hipLaunchKernelGGL((k_up9<T, 32, 8>), tiles(L.d, 32, 8), dim3(NTB), 0, s, L.d, opv(Acc<T>::op(L)),
                               (const T*)Acc<T>::vec(L).b, (const T*)Acc<T>::vec(L).x, Cl.d, Acc<T>::pr(L), e,
                               Acc<T>::vec(L).r, on);
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace

int mg_vcycle_dots(Ctx* c, double* partials, int max_blocks, int* dots_done) {
    Slots S{partials, max_blocks};
    if (c->prm.precision == SG_PREC_MIXED)
        return vcycle<float, float>(c, c->rf, c->zf, c->qf, S, dots_done, &c->sc->active);
    return vcycle<double, double>(c, c->r, c->z, c->q, S, dots_done, &c->sc->active);
}

namespace {
__global__ void __launch_bounds__(256) k_copy(size_t n, const double* __restrict__ a, double* __restrict__ b) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[i];
}
}  // namespace

// One launch of a fine-level kernel on the current data (measurement):
// 1 = down-sweep, 2 = up-sweep, 3 = plain copy of the level-0 padded vector.
int mg_launch_fine(Ctx* c, int which) {
    hipStream_t s = c->stream;
    const Level& L = c->lev[0];
    if (c->nlev < 2) return SG_ERR_STATE;
    const Level& Cl = c->lev[1];
    Slots S{c->red.partials + (size_t)7 * NRED * c->red.max_blocks, c->red.max_blocks};
    const bool mixed = c->prm.precision == SG_PREC_MIXED;
    if (which == 1) {
        if (mixed)
            launch_down5<float, float>(c->semi0 != 0, s, L.d, opv(L.Af), (const float*)c->rf, L.vf.x, Cl.d, L.Pf,
                                       Cl.vf.b, nullptr);
        else
            launch_down5<double, double>(c->semi0 != 0, s, L.d, opv(L.A), (const double*)c->r, L.v.x, Cl.d, L.P,
                                         Cl.v.b, nullptr);
    } else if (which == 2) {
        // the coarse correction sits in x when level 1 is the tail's first level, else in r
        const bool tail1 = c->nlev == 2 || (size_t)Cl.d.ni * Cl.d.nj <= (size_t)TAIL_MAX;
        if (mixed)
            launch_up5<float, float>(c->semi0 != 0, s, L.d, opv(L.Af), (const float*)c->rf, (const float*)L.vf.x,
                                     Cl.d, L.Pf, (const float*)(tail1 ? Cl.vf.x : Cl.vf.r), c->zf,
                                     (const float*)c->qf, S, nullptr);
        else
            launch_up5<double, double>(c->semi0 != 0, s, L.d, opv(L.A), (const double*)c->r, (const double*)L.v.x,
                                       Cl.d, L.P, (const double*)(tail1 ? Cl.v.x : Cl.v.r), c->z,
                                       (const double*)c->q, S, nullptr);
    } else if (which == 3) {
        const size_t n = L.d.alloc_elems();
        hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, (const double*)L.v.b,
                           L.v.r);
    } else {
        return SG_ERR_ARG;
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int mg_partial_blocks(Ctx* c) {
    const dim3 t = tiles(c->lev[0].d, GW, GH);
    return (int)(t.x * t.y);
}

}  // namespace sg
