// gft_reduce.hpp -- deterministic FP64 block reductions.
//
// Each block reduces per-thread partial sums (pairwise through wavefront
// shuffles, then LDS across waves) and writes one partial per quantity; a
// single-block finisher sums the per-block partials with Neumaier compensation
// in a fixed order.  The result does not depend on timing or atomics.
#pragma once

#include <hip/hip_runtime.h>

#include "gft_math.hpp"

namespace gft {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}

// Reduces K sums and M maxima over the block; thread 0 writes
// out[blockIdx.x * (K + M) + k].  lds must hold (blockDim.x / 64) * (K + M) doubles.
template <int K, int M>
__device__ __forceinline__ void block_reduce_store(const double* sums, const double* maxs, double* lds, double* out,
                                                   long long block_id) {
    const int lane = threadIdx.x % kWave, wave = threadIdx.x / kWave, nw = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double v = wave_sum(sums[k]);
        if (lane == 0) lds[wave * (K + M) + k] = v;
    }
#pragma unroll
    for (int m = 0; m < M; m++) {
        double v = wave_max(maxs[m]);
        if (lane == 0) lds[wave * (K + M) + K + m] = v;
    }
    __syncthreads();
    if (threadIdx.x < K + M) {
        const int q = threadIdx.x;
        double acc = lds[q];
        for (int w = 1; w < nw; w++) acc = (q < K) ? acc + lds[w * (K + M) + q] : fmax(acc, lds[w * (K + M) + q]);
        out[block_id * (K + M) + q] = acc;
    }
}

// One block of kFinThreads threads: out[q] = Neumaier sum (q < nsum) or max
// (q >= nsum) of partial[b * stride + q] over b = 0 .. nblocks-1 (stride >=
// nsum + nmax: a wider row is finished in several calls on column slices).  Thread
// t takes rows t, t + kFinThreads, ... for every q at once (independent
// accumulator chains), then thread q combines the per-thread (sum,
// compensation) pairs in thread order: a fixed order, so the result is
// deterministic.  nsum + nmax <= kFinWidth.
constexpr int kFinThreads = 256;
constexpr int kFinWidth = 12;
// Defined once, in gft_capi.hip (the library's other translation unit includes
// this header too).
__global__ void __launch_bounds__(kFinThreads) finalize_partials(const double* __restrict__ partial, long long nblocks,
                                                                 int stride, int nsum, int nmax,
                                                                 double* __restrict__ out);

}  // namespace gft
