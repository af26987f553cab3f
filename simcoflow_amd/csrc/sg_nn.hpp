// The learned warm start of the projection solve: a 4-level convolutional
// encoder-decoder (16/32/64/128 channels, 3x3 kernels, additive skips,
// ReLU, no biases) on a coarse level of the multigrid hierarchy.  BF16
// weights and activations, FP32 accumulation on the matrix cores.
#pragma once
#include "sg_internal.hpp"

namespace sg {
namespace nn {

typedef __bf16 bf16;

constexpr int NL = 15;   // convolution layers
constexpr int CIN0 = 8;  // input channels
struct LSpec {
    int cin, cout, lvl;
};
// encoder l0..l7, decoder l8..l13, output l14 (16 rows, only row 0 nonzero)
constexpr LSpec SPEC[NL] = {{8, 16, 0},   {16, 16, 0},  {16, 32, 1}, {32, 32, 1}, {32, 64, 2},
                            {64, 64, 2},  {64, 128, 3}, {128, 128, 3}, {128, 64, 2}, {64, 64, 2},
                            {64, 32, 1},  {32, 32, 1},  {32, 16, 0}, {16, 16, 0}, {16, 16, 0}};
__host__ __device__ constexpr int kpad(int c) { return (9 * c + 31) / 32 * 32; }

// activation tensor [B][H+2][W+2][C], zero one-pixel border
struct T4 {
    bf16* p = nullptr;
    int H = 0, W = 0, C = 0;
};
__host__ __device__ inline size_t t4off(const T4& t, int b, int y, int x) {
    return (((size_t)b * (t.H + 2) + y) * (t.W + 2) + x) * (size_t)t.C;
}

struct Net {
    int H = 0, W = 0, B = 0;  // level-0 size of the net, batch capacity
    size_t nw = 0;            // master weights (all layers)
    size_t woff[NL + 1] = {0};
    size_t foff[NL + 1] = {0};
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;  // master, gradient, Adam moments
    bf16 *wt = nullptr, *wf = nullptr;  // BF16 weights, flipped/transposed for the data gradient
    // forward activations
    T4 e0a, e0, d1, e1a, e1, d2, e2a, e2, d3, e3a, e3, u2, g2a, g2, u1, g1a, g1, u0, g0a, g0;
    // backward tensors
    T4 gout, P13, P12, du0, P11, P10, du1, P9, P8, du2, P7, P6, dd3, P5, P4, dd2, P3, P2, dd1, P1, P0;
    float* y = nullptr;     // output [B][H][W]
    float* loss = nullptr;  // device scalar
    int step = 0;
    hipStream_t stream = nullptr;
    std::vector<void*> allocs;
};

size_t num_weights();
int create(Net* n, int H, int W, int B, hipStream_t s);
void destroy(Net* n);
int set_weights(Net* n, const float* host);   // master <- host, refresh BF16 copies
int get_weights(Net* n, float* host);
// X: input [nb][H+2][W+2][8] (device, zero border); y <- output
int forward(Net* n, const bf16* X, int nb);
// loss = mean over pixels of mask*(y - t)^2, gradients into n->g
int backward(Net* n, const bf16* X, const float* t, const float* mask, int nb);
int adam_step(Net* n, float lr);

}  // namespace nn
}  // namespace sg
