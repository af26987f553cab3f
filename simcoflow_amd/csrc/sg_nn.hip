// Convolutional encoder-decoder of the learned warm start: forward,
// backward (data and weight gradients) and Adam, hand-written for the
// gfx950 matrix cores.  Convolutions are implicit GEMMs on
// v_mfma_f32_16x16x32_bf16: the output tile is 16 channels x 16 pixels, the
// reduction runs over (tap, input channel) in chunks of 32, activations are
// channel-innermost so one lane's 8 consecutive reduction elements are one
// 16-byte load.  Weight gradients reduce over pixels: 32-pixel tiles of the
// input and of the output gradient are staged transposed in LDS so the same
// instruction reads 8 consecutive pixels per lane.  No counterpart in the
// reference (it starts PCG from zero, ProjectionP_Mod.f90:319).
#include <cmath>

#include "sg_nn.hpp"

namespace sg {
namespace nn {

namespace {

typedef bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

enum Epi { E_RELU = 0, E_RELU_SKIP, E_OUT, E_MASK, E_PLAIN };

__device__ __forceinline__ floatx4 mfma(const bf16x8& a, const bf16x8& b, const floatx4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// out(b,y,x,co) = sum_{tap,ci} wt[co][tap*CIN+ci] * in(b, y+dy-1, x+dx-1, ci)
// One wave: 16*NPB pixels of one row and COUT/CG of the output channels
// (CG > 1 on small levels, where one wave per pixel group would leave most
// of the chip idle; each output still sums over (tap, ci) in the same
// order).  The epilogue
// applies ReLU (optionally after adding a skip tensor), masks by aux > 0
// (data gradient through a ReLU), writes plain, or writes channel 0 as FP32
// (the output layer).
template <int CIN, int COUT, int NPB, int CG>
__global__ void __launch_bounds__(256) k_conv(T4 in, const bf16* __restrict__ wt, T4 out, T4 aux,
                                              float* __restrict__ yo, int nb, int epi) {
    constexpr int KP = kpad(CIN);
    constexpr int MB = COUT / 16 / CG;
    static_assert(COUT % (16 * CG) == 0, "channel groups must split the 16-channel blocks");
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int tiles = in.W / (16 * NPB);
    const long wave_all = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave_all >= (long)nb * in.H * tiles * CG) return;
    const int g = (int)(wave_all % CG);
    const long wave = wave_all / CG;
    if (epi == E_OUT && g != 0) return;  // the output layer keeps channel 0 only
    const int tx = (int)(wave % tiles);
    const long t = wave / tiles;
    const int y = (int)(t % in.H), b = (int)(t / in.H);
    const int x0 = tx * 16 * NPB;
    floatx4 acc[MB][NPB];
#pragma unroll
    for (int mb = 0; mb < MB; ++mb)
#pragma unroll
        for (int q = 0; q < NPB; ++q) acc[mb][q] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bf16* wrow = wt + ((size_t)g * MB * 16 + r16) * KP + 8 * q4;
    for (int kc = 0; kc < KP; kc += 32) {
        const int k0 = kc + 8 * q4;
        const int tap = k0 / CIN, ci = k0 - tap * CIN;
        bf16x8 bv[NPB];
        if (tap < 9) {
            const int dy = tap / 3, dx = tap - 3 * dy;
            const bf16* src = in.p + t4off(in, b, y + dy, x0 + r16 + dx) + ci;
#pragma unroll
            for (int q = 0; q < NPB; ++q) bv[q] = *(const bf16x8*)(src + (size_t)q * 16 * CIN);
        } else {
#pragma unroll
            for (int q = 0; q < NPB; ++q)
#pragma unroll
                for (int e = 0; e < 8; ++e) bv[q][e] = (bf16)0.f;
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const bf16x8 av = *(const bf16x8*)(wrow + (size_t)mb * 16 * KP + kc);
#pragma unroll
            for (int q = 0; q < NPB; ++q) acc[mb][q] = mfma(av, bv[q], acc[mb][q]);
        }
    }
    // lane holds D[co = mb*16 + 4*q4 + r][px = x0 + q*16 + r16], r = 0..3
#pragma unroll
    for (int q = 0; q < NPB; ++q) {
        const int px = x0 + q * 16 + r16;
        if (epi == E_OUT) {
            if (q4 == 0) yo[((size_t)b * in.H + y) * in.W + px] = acc[0][q][0];
            continue;
        }
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
            const int co = (g * MB + mb) * 16 + 4 * q4;
            float v[4] = {acc[mb][q][0], acc[mb][q][1], acc[mb][q][2], acc[mb][q][3]};
            if (epi == E_RELU_SKIP || epi == E_MASK) {
                const bf16x4 a = *(const bf16x4*)(aux.p + t4off(aux, b, y + 1, px + 1) + co);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float av = (float)a[r];
                    v[r] = epi == E_MASK ? (av > 0.f ? v[r] : 0.f) : v[r] + av;
                }
            }
            if (epi == E_RELU || epi == E_RELU_SKIP) {
#pragma unroll
                for (int r = 0; r < 4; ++r) v[r] = v[r] > 0.f ? v[r] : 0.f;
            }
            bf16x4 o;
#pragma unroll
            for (int r = 0; r < 4; ++r) o[r] = (bf16)v[r];
            *(bf16x4*)(out.p + t4off(out, b, y + 1, px + 1) + co) = o;
        }
    }
}

// dW[co][tap*CIN+ci] += sum_px P(px, co) * X(px shifted by tap, ci)
// block: one tap (blockIdx.y), tiles_per_block tiles of 32 pixels in the
// flattened (b, y, x) order; partial sums are added to dW atomically.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256) k_dw(T4 X, T4 P, float* __restrict__ dW, int nb, int tiles_per_block) {
    constexpr int CI16 = (CIN + 15) / 16 * 16;
    constexpr int KP = kpad(CIN), MB = COUT / 16, NBK = CI16 / 16;
    constexpr int NBLK = MB * NBK;
    constexpr int PER = (NBLK + 3) / 4;
    constexpr int LDW = 40;  // 32 pixels + 8 (rows stay 16-byte aligned)
    __shared__ __attribute__((aligned(16))) bf16 gT[COUT * LDW];
    __shared__ __attribute__((aligned(16))) bf16 xT[CI16 * LDW];
    const int tap = blockIdx.y, dy = tap / 3, dx = tap - 3 * (tap / 3);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, r16 = lane & 15, q4 = lane >> 4;
    const int HW = P.H * P.W;
    const long npx = (long)nb * HW;
    floatx4 acc[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) acc[k] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int e = tid; e < (CI16 - CIN) * LDW; e += 256) xT[CIN * LDW + e] = (bf16)0.f;
    for (int tt = 0; tt < tiles_per_block; ++tt) {
        const long q0 = ((long)blockIdx.x * tiles_per_block + tt) * 32;
        if (q0 >= npx) break;  // uniform over the block
        __syncthreads();
        for (int e = tid; e < 32 * (COUT / 8); e += 256) {
            const int px = e / (COUT / 8), c8 = (e - px * (COUT / 8)) * 8;
            const long q = q0 + px;
            bf16x8 v;
            if (q < npx) {
                const int b = (int)(q / HW), rem = (int)(q - (long)b * HW);
                const int y = rem / P.W, x = rem - y * P.W;
                v = *(const bf16x8*)(P.p + t4off(P, b, y + 1, x + 1) + c8);
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = (bf16)0.f;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) gT[(c8 + k) * LDW + px] = v[k];
        }
        for (int e = tid; e < 32 * (CIN / 8); e += 256) {
            const int px = e / (CIN / 8), c8 = (e - px * (CIN / 8)) * 8;
            const long q = q0 + px;
            bf16x8 v;
            if (q < npx) {
                const int b = (int)(q / HW), rem = (int)(q - (long)b * HW);
                const int y = rem / P.W, x = rem - y * P.W;
                v = *(const bf16x8*)(X.p + t4off(X, b, y + dy, x + dx) + c8);
            } else {
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = (bf16)0.f;
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) xT[(c8 + k) * LDW + px] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int blk = wv + 4 * k;
            if (blk < NBLK) {
                const int mb = blk / NBK, nbk = blk - mb * NBK;
                const bf16x8 a = *(const bf16x8*)(gT + (mb * 16 + r16) * LDW + 8 * q4);
                const bf16x8 bb = *(const bf16x8*)(xT + (nbk * 16 + r16) * LDW + 8 * q4);
                acc[k] = mfma(a, bb, acc[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int blk = wv + 4 * k;
        if (blk >= NBLK) continue;
        const int mb = blk / NBK, nbk = blk - mb * NBK;
        const int ci = nbk * 16 + r16;
        if (ci >= CIN) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r)
            atomicAdd(dW + (size_t)(mb * 16 + 4 * q4 + r) * KP + tap * CIN + ci, acc[k][r]);
    }
}

// 2x2 average pooling, 8 channels per thread
__global__ void __launch_bounds__(256) k_pool(T4 in, T4 out, int nb) {
    const int c8n = out.C / 8;
    const long n = (long)nb * out.H * out.W * c8n;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int c8 = (int)(e % c8n) * 8;
    long t = e / c8n;
    const int x = (int)(t % out.W);
    t /= out.W;
    const int y = (int)(t % out.H), b = (int)(t / out.H);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) {
            const bf16x8 v = *(const bf16x8*)(in.p + t4off(in, b, 2 * y + a + 1, 2 * x + c + 1) + c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += (float)v[k];
        }
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (bf16)(0.25f * s[k]);
    *(bf16x8*)(out.p + t4off(out, b, y + 1, x + 1) + c8) = o;
}

// nearest-neighbour upsampling by 2
__global__ void __launch_bounds__(256) k_up(T4 in, T4 out, int nb) {
    const int c8n = out.C / 8;
    const long n = (long)nb * out.H * out.W * c8n;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int c8 = (int)(e % c8n) * 8;
    long t = e / c8n;
    const int x = (int)(t % out.W);
    t /= out.W;
    const int y = (int)(t % out.H), b = (int)(t / out.H);
    *(bf16x8*)(out.p + t4off(out, b, y + 1, x + 1) + c8) =
        *(const bf16x8*)(in.p + t4off(in, b, y / 2 + 1, x / 2 + 1) + c8);
}

// gradient through nearest upsampling and the ReLU of its input:
// P(y,x) = sum_{2x2} d(2y+a, 2x+c) * (act(y,x) > 0)
__global__ void __launch_bounds__(256) k_upT_mask(T4 d, T4 act, T4 P, int nb) {
    const int c8n = P.C / 8;
    const long n = (long)nb * P.H * P.W * c8n;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int c8 = (int)(e % c8n) * 8;
    long t = e / c8n;
    const int x = (int)(t % P.W);
    t /= P.W;
    const int y = (int)(t % P.H), b = (int)(t / P.H);
    float s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c) {
            const bf16x8 v = *(const bf16x8*)(d.p + t4off(d, b, 2 * y + a + 1, 2 * x + c + 1) + c8);
#pragma unroll
            for (int k = 0; k < 8; ++k) s[k] += (float)v[k];
        }
    const bf16x8 m = *(const bf16x8*)(act.p + t4off(act, b, y + 1, x + 1) + c8);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (bf16)((float)m[k] > 0.f ? s[k] : 0.f);
    *(bf16x8*)(P.p + t4off(P, b, y + 1, x + 1) + c8) = o;
}

// gradient at a tensor that feeds a skip and a 2x2 pooling, through its
// ReLU: P = (skip + dd(y/2,x/2)/4) * (act > 0)
__global__ void __launch_bounds__(256) k_poolT_add_mask(T4 dd, T4 skip, T4 act, T4 P, int nb) {
    const int c8n = P.C / 8;
    const long n = (long)nb * P.H * P.W * c8n;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const int c8 = (int)(e % c8n) * 8;
    long t = e / c8n;
    const int x = (int)(t % P.W);
    t /= P.W;
    const int y = (int)(t % P.H), b = (int)(t / P.H);
    const bf16x8 a = *(const bf16x8*)(dd.p + t4off(dd, b, y / 2 + 1, x / 2 + 1) + c8);
    const bf16x8 s = *(const bf16x8*)(skip.p + t4off(skip, b, y + 1, x + 1) + c8);
    const bf16x8 m = *(const bf16x8*)(act.p + t4off(act, b, y + 1, x + 1) + c8);
    bf16x8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = (bf16)((float)m[k] > 0.f ? (float)s[k] + 0.25f * (float)a[k] : 0.f);
    *(bf16x8*)(P.p + t4off(P, b, y + 1, x + 1) + c8) = o;
}

// loss = sum mask*(y-t)^2 / n ; gout channel 0 = 2 mask (y-t) / n
__global__ void __launch_bounds__(256) k_loss(const float* __restrict__ y, const float* __restrict__ t,
                                              const float* __restrict__ m, T4 g, int nb, float inv_n,
                                              float* __restrict__ loss) {
    __shared__ float red[4];
    const long n = (long)nb * g.H * g.W;
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    float l = 0.f;
    if (e < n) {
        const int x = (int)(e % g.W);
        const long r = e / g.W;
        const int yy = (int)(r % g.H), b = (int)(r / g.H);
        const float d = m[e] > 0.f ? y[e] - t[e] : 0.f;
        l = d * d * inv_n;
        g.p[t4off(g, b, yy + 1, x + 1)] = (bf16)(2.f * d * inv_n);
    }
    for (int o = 32; o >= 1; o >>= 1) l += __shfl_xor(l, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, red[0] + red[1] + red[2] + red[3]);
}

__global__ void __launch_bounds__(256) k_fill(float* __restrict__ a, size_t n, float v) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) a[e] = v;
}

__global__ void __launch_bounds__(256) k_to_bf16(const float* __restrict__ a, bf16* __restrict__ o, size_t n) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e < n) o[e] = (bf16)a[e];
}

// data-gradient weights of one layer: wf[ci][tap'*cout + co] =
// wt[co][(8-tap')*cin + ci], zero beyond 9*cout
__global__ void __launch_bounds__(256) k_flip(const bf16* __restrict__ wt, bf16* __restrict__ wf, int cin,
                                              int cout) {
    const int kpo = kpad(cout), kpi = kpad(cin);
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= cin * kpo) return;
    const int ci = e / kpo, k = e - ci * kpo;
    bf16 v = (bf16)0.f;
    if (k < 9 * cout) {
        const int tp = k / cout, co = k - tp * cout;
        v = wt[(size_t)co * kpi + (8 - tp) * cin + ci];
    }
    wf[e] = v;
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, size_t n, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.f - b1) * gg;
    const float vv = b2 * v[e] + (1.f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm * c1) / (sqrtf(vv * c2) + eps);
}

inline unsigned blocks(long n) { return (unsigned)((n + 255) / 256); }

template <int CIN, int COUT, int NPB, int CG>
void conv_launch(Net* n, T4 in, const bf16* w, T4 out, T4 aux, float* yo, int nb, int epi, long waves) {
    hipLaunchKernelGGL((k_conv<CIN, COUT, NPB, CG>), dim3((unsigned)((waves * CG + 3) / 4)), dim3(256), 0, n->stream,
                       in, w, out, aux, yo, nb, epi);
}

template <int CIN, int COUT, int NPB>
void conv_split(Net* n, T4 in, const bf16* w, T4 out, T4 aux, float* yo, int nb, int epi, long waves) {
    // split the output channels while there are fewer than ~4 waves per CU
    constexpr int MB = COUT / 16;
    int cg = 1;
    while (cg < MB && waves * cg < 1024) cg *= 2;
    if (cg >= 8 && MB % 8 == 0)
        conv_launch<CIN, COUT, NPB, (MB % 8 == 0 ? 8 : 1)>(n, in, w, out, aux, yo, nb, epi, waves);
    else if (cg >= 4 && MB % 4 == 0)
        conv_launch<CIN, COUT, NPB, (MB % 4 == 0 ? 4 : 1)>(n, in, w, out, aux, yo, nb, epi, waves);
    else if (cg >= 2 && MB % 2 == 0)
        conv_launch<CIN, COUT, NPB, (MB % 2 == 0 ? 2 : 1)>(n, in, w, out, aux, yo, nb, epi, waves);
    else
        conv_launch<CIN, COUT, NPB, 1>(n, in, w, out, aux, yo, nb, epi, waves);
}

template <int CIN, int COUT>
int conv(Net* n, T4 in, const bf16* w, T4 out, T4 aux, float* yo, int nb, int epi) {
    // four 16-pixel groups per wave where that still leaves ~4 waves per CU
    if (in.W % 64 == 0 && (long)nb * in.H * (in.W / 64) >= 1024)
        conv_split<CIN, COUT, 4>(n, in, w, out, aux, yo, nb, epi, (long)nb * in.H * (in.W / 64));
    else
        conv_split<CIN, COUT, 1>(n, in, w, out, aux, yo, nb, epi, (long)nb * in.H * (in.W / 16));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

template <int CIN, int COUT>
int dw(Net* n, int l, T4 X, T4 P, int nb) {
    const long tiles = ((long)nb * P.H * P.W + 31) / 32;
    int per = (int)((tiles + 255) / 256);
    if (per < 1) per = 1;
    const unsigned gx = (unsigned)((tiles + per - 1) / per);
    hipLaunchKernelGGL((k_dw<CIN, COUT>), dim3(gx, 9), dim3(256), 0, n->stream, X, P, n->g + n->woff[l], nb, per);
    SG_HIP(hipGetLastError());
    return SG_OK;
}

#define NN_TRY(x)               \
    do {                        \
        int _r = (x);           \
        if (_r) return _r;      \
    } while (0)

const bf16* WT(Net* n, int l) { return n->wt + n->woff[l]; }
const bf16* WF(Net* n, int l) { return n->wf + n->foff[l]; }

int refresh_bf16(Net* n) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks((long)n->nw)), dim3(256), 0, n->stream, (const float*)n->w, n->wt,
                       n->nw);
    for (int l = 1; l < NL; ++l) {
        const int cin = SPEC[l].cin, cout = SPEC[l].cout;
        hipLaunchKernelGGL(k_flip, dim3(blocks((long)cin * kpad(cout))), dim3(256), 0, n->stream,
                           (const bf16*)(n->wt + n->woff[l]), n->wf + n->foff[l], cin, cout);
    }
    SG_HIP(hipGetLastError());
    return SG_OK;
}

}  // namespace

size_t num_weights() {
    size_t s = 0;
    for (int l = 0; l < NL; ++l) s += (size_t)SPEC[l].cout * kpad(SPEC[l].cin);
    return s;
}

void destroy(Net* n) {
    for (void* p : n->allocs) (void)hipFree(p);
    n->allocs.clear();
}

int create(Net* n, int H, int W, int B, hipStream_t s) {
    if (H % 8 || W % 128 || H <= 0 || B <= 0) {
        set_error("learned warm start: the net needs H % 8 == 0 and W % 128 == 0");
        return SG_ERR_ARG;
    }
    n->H = H;
    n->W = W;
    n->B = B;
    n->stream = s;
    size_t o = 0, f = 0;
    for (int l = 0; l < NL; ++l) {
        n->woff[l] = o;
        n->foff[l] = f;
        o += (size_t)SPEC[l].cout * kpad(SPEC[l].cin);
        f += (size_t)SPEC[l].cin * kpad(SPEC[l].cout);
    }
    n->woff[NL] = o;
    n->foff[NL] = f;
    n->nw = o;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        if (hipMemset(p, 0, bytes) != hipSuccess) {
            (void)hipFree(p);
            return nullptr;
        }
        n->allocs.push_back(p);
        return p;
    };
    n->w = (float*)alloc(o * 4);
    n->g = (float*)alloc(o * 4);
    n->m = (float*)alloc(o * 4);
    n->v = (float*)alloc(o * 4);
    n->wt = (bf16*)alloc(o * 2);
    n->wf = (bf16*)alloc(f * 2);
    n->y = (float*)alloc((size_t)B * H * W * 4);
    n->loss = (float*)alloc(4);
    if (!n->w || !n->g || !n->m || !n->v || !n->wt || !n->wf || !n->y || !n->loss) {
        destroy(n);
        return SG_ERR_ALLOC;
    }
    bool ok = true;
    auto t4 = [&](int lvl, int C) {
        T4 t;
        t.H = H >> lvl;
        t.W = W >> lvl;
        t.C = C;
        t.p = (bf16*)alloc((size_t)B * (t.H + 2) * (t.W + 2) * C * 2);
        if (!t.p) ok = false;
        return t;
    };
    n->e0a = t4(0, 16); n->e0 = t4(0, 16); n->d1 = t4(1, 16);
    n->e1a = t4(1, 32); n->e1 = t4(1, 32); n->d2 = t4(2, 32);
    n->e2a = t4(2, 64); n->e2 = t4(2, 64); n->d3 = t4(3, 64);
    n->e3a = t4(3, 128); n->e3 = t4(3, 128);
    n->u2 = t4(2, 128); n->g2a = t4(2, 64); n->g2 = t4(2, 64);
    n->u1 = t4(1, 64); n->g1a = t4(1, 32); n->g1 = t4(1, 32);
    n->u0 = t4(0, 32); n->g0a = t4(0, 16); n->g0 = t4(0, 16);
    n->gout = t4(0, 16); n->P13 = t4(0, 16); n->P12 = t4(0, 16); n->du0 = t4(0, 32);
    n->P11 = t4(1, 32); n->P10 = t4(1, 32); n->du1 = t4(1, 64);
    n->P9 = t4(2, 64); n->P8 = t4(2, 64); n->du2 = t4(2, 128);
    n->P7 = t4(3, 128); n->P6 = t4(3, 128); n->dd3 = t4(3, 64);
    n->P5 = t4(2, 64); n->P4 = t4(2, 64); n->dd2 = t4(2, 32);
    n->P3 = t4(1, 32); n->P2 = t4(1, 32); n->dd1 = t4(1, 16);
    n->P1 = t4(0, 16); n->P0 = t4(0, 16);
    if (!ok) {
        destroy(n);
        return SG_ERR_ALLOC;
    }
    return SG_OK;
}

int set_weights(Net* n, const float* host) {
    SG_HIP(hipMemcpyAsync(n->w, host, n->nw * 4, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_fill, dim3(blocks((long)n->nw)), dim3(256), 0, n->stream, n->m, n->nw, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(blocks((long)n->nw)), dim3(256), 0, n->stream, n->v, n->nw, 0.f);
    n->step = 0;
    NN_TRY(refresh_bf16(n));
    SG_HIP(hipStreamSynchronize(n->stream));
    return SG_OK;
}

int get_weights(Net* n, float* host) {
    SG_HIP(hipMemcpyAsync(host, n->w, n->nw * 4, hipMemcpyDeviceToHost, n->stream));
    SG_HIP(hipStreamSynchronize(n->stream));
    return SG_OK;
}

int forward(Net* n, const bf16* Xp, int nb) {
    if (nb < 1 || nb > n->B) {
        set_error("learned warm start: batch larger than the net's capacity");
        return SG_ERR_ARG;
    }
    T4 X;
    X.p = const_cast<bf16*>(Xp);
    X.H = n->H;
    X.W = n->W;
    X.C = CIN0;
    const T4 none;
    const hipStream_t s = n->stream;
    auto pool = [&](T4 a, T4 b) {
        hipLaunchKernelGGL(k_pool, dim3(blocks((long)nb * b.H * b.W * (b.C / 8))), dim3(256), 0, s, a, b, nb);
    };
    auto up = [&](T4 a, T4 b) {
        hipLaunchKernelGGL(k_up, dim3(blocks((long)nb * b.H * b.W * (b.C / 8))), dim3(256), 0, s, a, b, nb);
    };
    NN_TRY((conv<8, 16>(n, X, WT(n, 0), n->e0a, none, nullptr, nb, E_RELU)));
    NN_TRY((conv<16, 16>(n, n->e0a, WT(n, 1), n->e0, none, nullptr, nb, E_RELU)));
    pool(n->e0, n->d1);
    NN_TRY((conv<16, 32>(n, n->d1, WT(n, 2), n->e1a, none, nullptr, nb, E_RELU)));
    NN_TRY((conv<32, 32>(n, n->e1a, WT(n, 3), n->e1, none, nullptr, nb, E_RELU)));
    pool(n->e1, n->d2);
    NN_TRY((conv<32, 64>(n, n->d2, WT(n, 4), n->e2a, none, nullptr, nb, E_RELU)));
    NN_TRY((conv<64, 64>(n, n->e2a, WT(n, 5), n->e2, none, nullptr, nb, E_RELU)));
    pool(n->e2, n->d3);
    NN_TRY((conv<64, 128>(n, n->d3, WT(n, 6), n->e3a, none, nullptr, nb, E_RELU)));
    NN_TRY((conv<128, 128>(n, n->e3a, WT(n, 7), n->e3, none, nullptr, nb, E_RELU)));
    up(n->e3, n->u2);
    NN_TRY((conv<128, 64>(n, n->u2, WT(n, 8), n->g2a, n->e2, nullptr, nb, E_RELU_SKIP)));
    NN_TRY((conv<64, 64>(n, n->g2a, WT(n, 9), n->g2, none, nullptr, nb, E_RELU)));
    up(n->g2, n->u1);
    NN_TRY((conv<64, 32>(n, n->u1, WT(n, 10), n->g1a, n->e1, nullptr, nb, E_RELU_SKIP)));
    NN_TRY((conv<32, 32>(n, n->g1a, WT(n, 11), n->g1, none, nullptr, nb, E_RELU)));
    up(n->g1, n->u0);
    NN_TRY((conv<32, 16>(n, n->u0, WT(n, 12), n->g0a, n->e0, nullptr, nb, E_RELU_SKIP)));
    NN_TRY((conv<16, 16>(n, n->g0a, WT(n, 13), n->g0, none, nullptr, nb, E_RELU)));
    NN_TRY((conv<16, 16>(n, n->g0, WT(n, 14), none, none, n->y, nb, E_OUT)));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int backward(Net* n, const bf16* Xp, const float* t, const float* mask, int nb) {
    NN_TRY(forward(n, Xp, nb));
    T4 X;
    X.p = const_cast<bf16*>(Xp);
    X.H = n->H;
    X.W = n->W;
    X.C = CIN0;
    const T4 none;
    const hipStream_t s = n->stream;
    hipLaunchKernelGGL(k_fill, dim3(blocks((long)n->nw)), dim3(256), 0, s, n->g, n->nw, 0.f);
    hipLaunchKernelGGL(k_fill, dim3(1), dim3(256), 0, s, n->loss, (size_t)1, 0.f);
    const long npx = (long)nb * n->H * n->W;
    hipLaunchKernelGGL(k_loss, dim3(blocks(npx)), dim3(256), 0, s, (const float*)n->y, t, mask, n->gout, nb,
                       1.f / (float)npx, n->loss);
    auto upT = [&](T4 d, T4 act, T4 P) {
        hipLaunchKernelGGL(k_upT_mask, dim3(blocks((long)nb * P.H * P.W * (P.C / 8))), dim3(256), 0, s, d, act, P,
                           nb);
    };
    auto poolT = [&](T4 dd, T4 skip, T4 act, T4 P) {
        hipLaunchKernelGGL(k_poolT_add_mask, dim3(blocks((long)nb * P.H * P.W * (P.C / 8))), dim3(256), 0, s, dd,
                           skip, act, P, nb);
    };
    NN_TRY((dw<16, 16>(n, 14, n->g0, n->gout, nb)));
    NN_TRY((conv<16, 16>(n, n->gout, WF(n, 14), n->P13, n->g0, nullptr, nb, E_MASK)));
    NN_TRY((dw<16, 16>(n, 13, n->g0a, n->P13, nb)));
    NN_TRY((conv<16, 16>(n, n->P13, WF(n, 13), n->P12, n->g0a, nullptr, nb, E_MASK)));
    NN_TRY((dw<32, 16>(n, 12, n->u0, n->P12, nb)));
    NN_TRY((conv<16, 32>(n, n->P12, WF(n, 12), n->du0, none, nullptr, nb, E_PLAIN)));
    upT(n->du0, n->g1, n->P11);
    NN_TRY((dw<32, 32>(n, 11, n->g1a, n->P11, nb)));
    NN_TRY((conv<32, 32>(n, n->P11, WF(n, 11), n->P10, n->g1a, nullptr, nb, E_MASK)));
    NN_TRY((dw<64, 32>(n, 10, n->u1, n->P10, nb)));
    NN_TRY((conv<32, 64>(n, n->P10, WF(n, 10), n->du1, none, nullptr, nb, E_PLAIN)));
    upT(n->du1, n->g2, n->P9);
    NN_TRY((dw<64, 64>(n, 9, n->g2a, n->P9, nb)));
    NN_TRY((conv<64, 64>(n, n->P9, WF(n, 9), n->P8, n->g2a, nullptr, nb, E_MASK)));
    NN_TRY((dw<128, 64>(n, 8, n->u2, n->P8, nb)));
    NN_TRY((conv<64, 128>(n, n->P8, WF(n, 8), n->du2, none, nullptr, nb, E_PLAIN)));
    upT(n->du2, n->e3, n->P7);
    NN_TRY((dw<128, 128>(n, 7, n->e3a, n->P7, nb)));
    NN_TRY((conv<128, 128>(n, n->P7, WF(n, 7), n->P6, n->e3a, nullptr, nb, E_MASK)));
    NN_TRY((dw<64, 128>(n, 6, n->d3, n->P6, nb)));
    NN_TRY((conv<128, 64>(n, n->P6, WF(n, 6), n->dd3, none, nullptr, nb, E_PLAIN)));
    poolT(n->dd3, n->P8, n->e2, n->P5);
    NN_TRY((dw<64, 64>(n, 5, n->e2a, n->P5, nb)));
    NN_TRY((conv<64, 64>(n, n->P5, WF(n, 5), n->P4, n->e2a, nullptr, nb, E_MASK)));
    NN_TRY((dw<32, 64>(n, 4, n->d2, n->P4, nb)));
    NN_TRY((conv<64, 32>(n, n->P4, WF(n, 4), n->dd2, none, nullptr, nb, E_PLAIN)));
    poolT(n->dd2, n->P10, n->e1, n->P3);
    NN_TRY((dw<32, 32>(n, 3, n->e1a, n->P3, nb)));
    NN_TRY((conv<32, 32>(n, n->P3, WF(n, 3), n->P2, n->e1a, nullptr, nb, E_MASK)));
    NN_TRY((dw<16, 32>(n, 2, n->d1, n->P2, nb)));
    NN_TRY((conv<32, 16>(n, n->P2, WF(n, 2), n->dd1, none, nullptr, nb, E_PLAIN)));
    poolT(n->dd1, n->P12, n->e0, n->P1);
    NN_TRY((dw<16, 16>(n, 1, n->e0a, n->P1, nb)));
    NN_TRY((conv<16, 16>(n, n->P1, WF(n, 1), n->P0, n->e0a, nullptr, nb, E_MASK)));
    NN_TRY((dw<8, 16>(n, 0, X, n->P0, nb)));
    SG_HIP(hipGetLastError());
    return SG_OK;
}

int adam_step(Net* n, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    n->step += 1;
    const float c1 = 1.f / (1.f - powf(b1, (float)n->step));
    const float c2 = 1.f / (1.f - powf(b2, (float)n->step));
    hipLaunchKernelGGL(k_adam, dim3(blocks((long)n->nw)), dim3(256), 0, n->stream, n->w, (const float*)n->g, n->m,
                       n->v, n->nw, lr, b1, b2, c1, c2, eps);
    SG_HIP(hipGetLastError());
    return refresh_bf16(n);
}

}  // namespace nn
}  // namespace sg
