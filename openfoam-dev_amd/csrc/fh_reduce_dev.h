// Deterministic block reductions shared by the reduction kernels (included
// by the kernel sources only): each block leaves one partial per summed
// value, in a fixed order, for a finalize that adds them in a fixed order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fh
{

#ifndef FH_BLOCK
#define FH_BLOCK 256
#endif

__device__ __forceinline__ int32_t sell_base2(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

// deterministic two-value block sum; results valid in thread 0
__device__ __forceinline__ void block_sum2(double& a, double& b, double* smem)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
    {
        smem[wid] = a;
        smem[FH_BLOCK/64 + wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < FH_BLOCK/64; i++) { sa += smem[i]; sb += smem[FH_BLOCK/64 + i]; }
        a = sa;
        b = sb;
    }
}

__device__ __forceinline__ void write_partials(double a, double b, double* partials,
                                               int32_t nBlocks)
{
    __shared__ double smem[2*FH_BLOCK/64];
    block_sum2(a, b, smem);
    if (threadIdx.x == 0)
    {
        partials[blockIdx.x] = a;
        partials[nBlocks + blockIdx.x] = b;
    }
}

} // namespace fh
