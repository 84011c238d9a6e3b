// Deterministic grid reductions without atomics or device-scope fences: every reducing
// kernel runs RG blocks of RBS threads over contiguous chunks and writes one partial per
// block; the consumer kernel sums the RG partials in a fixed order in every block, so all
// blocks see bit-identical scalars and take identical decisions.
#pragma once
#include "cg_common.hpp"

namespace cg {

constexpr int RBS = 256;   // reducing kernels (measured best with RG = 1024 on 2.1 M cells)
constexpr int RG = 1024;   // blocks of every reducing kernel

template <int NT, int NV> CG_DEV void block_sum(double (&v)[NV])
{
    constexpr int NW = NT / 64;
    __shared__ double sm[NV][NW];
#pragma unroll
    for (int k = 0; k < NV; ++k)
        for (int off = 32; off > 0; off >>= 1) v[k] += __shfl_down(v[k], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) sm[k][w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double t = 0.0;
        for (int q = 0; q < NW; ++q) t += sm[k][q];
        v[k] = t;
    }
    __syncthreads();
}

// fixed-order sums of the RG partials of NV slots: the same values in every block
template <int NV>
CG_DEV void sum_slots(const double* __restrict__ part, const int (&slot)[NV], double (&v)[NV])
{
    static_assert(RG % RBS == 0, "partials must split evenly over the block");
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        v[k] = 0.0;
#pragma unroll
        for (int q = 0; q < RG / RBS; ++q) v[k] += part[slot[k] * RG + q * RBS + threadIdx.x];
    }
    block_sum<RBS, NV>(v);
}

} // namespace cg
