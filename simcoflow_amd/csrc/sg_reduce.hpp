// Deterministic FP64 reductions for gfx950: 64-lane wavefront butterfly
// (__shfl_xor, lowered to DPP/ds_swizzle), then across the block's
// wavefronts through LDS, then across blocks by a last-block-done ticket.
// The last block adds the per-block partials in block order, so a sum does
// not depend on the order in which blocks finish.
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

// Grid-wide sum.  Every block calls it once with its per-thread values; the
// block that finishes last writes the sums to out[0..NV).  Returns true in
// the threads of that last block (thread 0 holds nothing special).
template <int NV, int NT>
__device__ inline bool grid_reduce(double (&v)[NV], double* partials, int max_blocks,
                                   unsigned int* ticket, double* out) {
    __shared__ bool is_last;
    block_reduce<NV, NT>(v);
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    const int nblocks = gridDim.x * gridDim.y;
    const int bid = blockIdx.x + blockIdx.y * gridDim.x;
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[(size_t)k * max_blocks + bid] = v[k];
        __threadfence();
        unsigned int t = atomicAdd(ticket, 1u);
        is_last = (t == (unsigned int)(nblocks - 1));
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) s[k] = 0.0;
    for (int b = tid; b < nblocks; b += NT) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            s[k] += partials[(size_t)k * max_blocks + b];
        }
    }
    block_reduce<NV, NT>(s);
    if (tid == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) out[k] = s[k];
        *ticket = 0u;
        __threadfence();
    }
    return true;
}

}  // namespace sg
