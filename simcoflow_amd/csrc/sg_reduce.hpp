// Deterministic FP64 reductions for gfx950 in two phases:
//   1. every block reduces its values -- 64-lane wavefront butterfly
//      (__shfl_xor, lowered to DPP/ds_swizzle), then across the block's
//      wavefronts through LDS -- and stores one partial per block with a
//      plain store;
//   2. a one-block finalize kernel adds the partials in block order.
// The kernel boundary between the phases makes the partials visible to the
// finalize block without any device-scope fence inside the producer (a fence
// per block forces an L2 write-back on gfx950 and costs far more than the
// extra launch).  Sums do not depend on the order in which blocks finish.
#pragma once
#include "sg_internal.hpp"

namespace sg {

template <int NV>
__device__ inline void wave_reduce(double (&v)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v[k] += __shfl_xor(v[k], m, 64);
    }
}

// Reduce NV values over the whole block; the result is valid in thread 0.
template <int NV, int NT>
__device__ inline void block_reduce(double (&v)[NV]) {
    __shared__ double lds[NT / 64][NV];
    wave_reduce<NV>(v);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int lane = tid & 63, wid = tid >> 6;
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) lds[wid][k] = v[k];
    }
    __syncthreads();
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int w = 0; w < NT / 64; ++w) s += lds[w][k];
            v[k] = s;
        }
    }
    __syncthreads();
}

// Phase 1: store this block's sums at partials[k*max_blocks + block].
template <int NV, int NT>
__device__ inline void block_partials(double (&v)[NV], double* partials, int max_blocks) {
    block_reduce<NV, NT>(v);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (tid == 0) {
        const int bid = blockIdx.x + blockIdx.y * gridDim.x;
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[(size_t)k * max_blocks + bid] = v[k];
    }
}

// Phase 2 (inside a one-block kernel of NT threads): sums of the first
// nblocks partials of each of the NV planes; valid in thread 0.
template <int NV, int NT>
__device__ inline void finalize_sum(const double* partials, int max_blocks, int nblocks, double (&s)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += NT) {
#pragma unroll
        for (int k = 0; k < NV; ++k) s[k] += partials[(size_t)k * max_blocks + b];
    }
    block_reduce<NV, NT>(s);
}

}  // namespace sg
