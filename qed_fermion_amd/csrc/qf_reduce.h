// qf_reduce.h -- deterministic FP64 block reductions.
//
// Every reduction in this library has a fixed summation tree (lane-strided
// partial sums, then xor-shuffle butterflies, then waves in index order), so
// the same input gives bit-identical sums in every block and on every replay.
#pragma once
#include "qf_common.h"

namespace qf {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum over the block of up to NV values per thread.  red: LDS scratch of
// NV * 16 doubles.  Result valid in every thread.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* red) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * 16 + w] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
        for (int i = 0; i < nw; ++i) s += red[k * 16 + i];
        v[k] = s;
    }
    __syncthreads();
}

// Sum n doubles (fixed order) using the first wave; result broadcast through
// LDS to the whole block.  Called by all threads.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* const (&src)[NV], int n, double (&out)[NV], double* red) {
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = 0.0;
            for (int i = lane; i < n; i += 64) s += src[k][i];
            s = wave_sum(s);
            if (lane == 0) red[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = red[k];
    __syncthreads();
}

// sum_partials split in two so that the loads can be issued early: the first
// NP loads of each lane of the first wave are made by load_partials (before
// the caller's other loads), the rest and the sum by sum_partials_pf.  Same
// summation order as sum_partials; the sum starts from the first partial
// instead of 0.0 + it (equal except for the sign of a zero).  The loads are
// unconditional (clamped index, masked at the sum) so that no wait is placed
// right behind them.
template <int NV, int NP>
struct Partials { double v[NV][NP]; };

template <int NV, int NP>
__device__ __forceinline__ Partials<NV, NP> load_partials(const double* const (&src)[NV], int n) {
    const int lane = threadIdx.x & 63;
    Partials<NV, NP> p{};
    if (threadIdx.x < 64 && n > 0) {   // uniform per wave
#pragma unroll
        for (int k = 0; k < NV; ++k)
#pragma unroll
            for (int j = 0; j < NP; ++j) p.v[k][j] = src[k][min(lane + 64 * j, n - 1)];
    }
    return p;
}

template <int NV, int NP>
__device__ __forceinline__ void sum_partials_pf(const Partials<NV, NP>& p, const double* const (&src)[NV], int n,
                                                double (&out)[NV], double* red) {
    const int lane = threadIdx.x & 63;
    if (threadIdx.x < 64) {
#pragma unroll
        for (int k = 0; k < NV; ++k) {
            double s = lane < n ? p.v[k][0] : 0.0;
#pragma unroll
            for (int j = 1; j < NP; ++j)
                if (lane + 64 * j < n) s += p.v[k][j];
            for (int i = lane + 64 * NP; i < n; i += 64) s += src[k][i];
            s = wave_sum(s);
            if (lane == 0) red[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) out[k] = red[k];
    __syncthreads();
}

}  // namespace qf
