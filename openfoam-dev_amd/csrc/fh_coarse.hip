// Coarse correction of the tile-local preconditioner (mixed PCG on boxes).
//
// The tile-local DIC drops the couplings between 16^3 tiles; its iteration
// count grows with the number of tiles per direction.  The two-level form
// adds a correction on the space of piecewise-constant tile functions P:
//     M^-1 r = M_tile^-1 r + P Ac^-1 P^T r,   Ac = P^T A P,
// symmetric positive (negative) definite like A, so PCG applies unchanged.
// Ac (one row per tile: the tile's diagonal sum plus twice its inner
// couplings, and the couplings to its six neighbour tiles) is assembled on
// the device in a fixed summation order, factored and inverted once per
// coefficient update (rocSOLVER Cholesky, FP64), and kept as a dense FP32
// inverse; each apply is one matrix-vector product over it.
#include "fh_kernels.h"
#include "fh_solver_kernels.h"
#include "fh_finalize_dev.h"
#include "fh_coarse.h"

namespace fh
{

#define CO_B 256

// one block per 16^3 tile: the tile's diagonal sum and inner couplings
// (counted twice), and its couplings across its lower x, y and z faces;
// each thread takes 16 cells of the tile in a fixed order, then a fixed
// block reduction
__global__ __launch_bounds__(CO_B)
void k_coarse_tiles(int32_t nx, int32_t ny, int32_t nz, int32_t nTI, int32_t nTJ,
                    const double* __restrict__ diag, const double* __restrict__ lK,
                    const double* __restrict__ lJ, const double* __restrict__ lI,
                    double* __restrict__ ctile)
{
    __shared__ double red[4][CO_B/64];
    const int32_t t = blockIdx.x;
    const int32_t TI = t % nTI, TJ = (t/nTI) % nTJ, TK = t/(nTI*nTJ);
    const int64_t sy = nx, sz = (int64_t)nx*ny;
    double dsum = 0.0, inner = 0.0, cx = 0.0, cy = 0.0, cz = 0.0;
    for (int m = 0; m < 16; m++)
    {
        const int32_t loc = threadIdx.x + CO_B*m;          // 0 .. 4095
        const int32_t x = TI*16 + (loc & 15), y = TJ*16 + ((loc >> 4) & 15);
        const int32_t z = TK*16 + (loc >> 8);
        if (x >= nx || y >= ny || z >= nz) continue;
        const int64_t c = x + sy*y + sz*z;
        dsum += diag[c];
        if (x > 0) { if (x & 15) inner += lI[c]; else cx += lI[c]; }
        if (y > 0) { if (y & 15) inner += lJ[c]; else cy += lJ[c]; }
        if (z > 0) { if (z & 15) inner += lK[c]; else cz += lK[c]; }
    }
    double v[4] = {dsum + 2.0*inner, cx, cy, cz};
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
        double a = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
        if (lane == 0) red[q][wid] = a;
    }
    __syncthreads();
    if (threadIdx.x < 4)
    {
        double a = 0.0;
        for (int w = 0; w < CO_B/64; w++) a += red[threadIdx.x][w];
        ctile[4*(int64_t)t + threadIdx.x] = a;
    }
}

// the sign that makes the coarse matrix positive definite: that of its
// first diagonal entry (the Laplacian is negative definite)
__device__ __forceinline__ double coarse_sign(const double* ctile)
{
    return ctile[0] < 0.0 ? -1.0 : 1.0;
}

// the dense coarse matrix times coarse_sign, column-major
__global__ void k_coarse_dense(int32_t nT, int32_t nTI, int32_t nTJ, const double* __restrict__ ctile,
                               double* __restrict__ dense)
{
    const int64_t idx = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nT*nT) return;
    const double sign = coarse_sign(ctile);
    const int32_t i = (int32_t)(idx % nT), j = (int32_t)(idx/nT);
    double v = 0.0;
    const int32_t lo = i < j ? i : j, hi = i < j ? j : i, d = hi - lo;
    const int32_t TIh = hi % nTI, TJh = (hi/nTI) % nTJ;
    if (d == 0) v = ctile[4*(int64_t)i];
    else if (d == 1 && TIh > 0) v = ctile[4*(int64_t)hi + 1];
    else if (d == nTI && TJh > 0) v = ctile[4*(int64_t)hi + 2];
    else if (d == nTI*nTJ) v = ctile[4*(int64_t)hi + 3];
    dense[idx] = sign*v;
}

// the FP32 inverse of Ac from the lower triangle of inv(sign*Ac): full
// symmetric, rows of ld >= nT floats (ld a multiple of 4, zero padding)
__global__ void k_coarse_inv32(int32_t nT, int32_t ld, const double* __restrict__ dense,
                               const double* __restrict__ ctile, float* __restrict__ inv)
{
    const int64_t idx = blockIdx.x*(int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)nT*ld) return;
    const double sign = coarse_sign(ctile);
    const int32_t j = (int32_t)(idx % ld), i = (int32_t)(idx/ld);
    float v = 0.f;
    if (j < nT)
        v = (float)(sign*(i >= j ? dense[i + (int64_t)nT*j] : dense[j + (int64_t)nT*i]));
    inv[idx] = v;
}

void launch_coarse_tiles(int32_t nx, int32_t ny, int32_t nz, int32_t nTI, int32_t nTJ,
                         int32_t nT, const double* diag, const double* const dir[3],
                         double* ctile, hipStream_t s)
{
    hipLaunchKernelGGL(k_coarse_tiles, dim3(nT), dim3(CO_B), 0, s, nx, ny, nz, nTI, nTJ, diag,
                       dir[0], dir[1], dir[2], ctile);
}

void launch_coarse_dense(int32_t nT, int32_t nTI, int32_t nTJ, const double* ctile,
                         double* dense, hipStream_t s)
{
    const int64_t n = (int64_t)nT*nT;
    hipLaunchKernelGGL(k_coarse_dense, dim3((unsigned)((n + CO_B - 1)/CO_B)), dim3(CO_B), 0, s,
                       nT, nTI, nTJ, ctile, dense);
}

void launch_coarse_inv32(int32_t nT, int32_t ld, const double* dense, const double* ctile,
                         float* inv, hipStream_t s)
{
    const int64_t n = (int64_t)nT*ld;
    hipLaunchKernelGGL(k_coarse_inv32, dim3((unsigned)((n + CO_B - 1)/CO_B)), dim3(CO_B), 0, s,
                       nT, ld, dense, ctile, inv);
}

// z = Ainv rc (one wavefront per row, four rows per block) and sum z*rc;
// block 0 also adds the fine partials of the tile-local apply (sum w*r, nF
// of them), so that the last block finalizes the whole w.r of the
// two-level apply (op FIN_PCG_WARA)
__global__ __launch_bounds__(CO_B)
void k_coarse_apply(int32_t nT, int32_t ld, const float* __restrict__ inv,
                    const double* __restrict__ rc, float* __restrict__ z,
                    const double* __restrict__ fine, int32_t nF, double* partials,
                    FusedFinalize F, const DevState* st)
{
    if (st && st->done) return;
    extern __shared__ float rcs[];
    for (int32_t j = threadIdx.x; j < ld; j += CO_B) rcs[j] = j < nT ? (float)rc[j] : 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t row = blockIdx.x*4 + wid;
    double part = 0.0;
    if (row < nT)
    {
        const float* a = inv + (int64_t)row*ld;
        float acc = 0.f;
        for (int32_t j = lane*4; j < ld; j += 256)
        {
            const float4 v = *reinterpret_cast<const float4*>(a + j);
            acc += v.x*rcs[j] + v.y*rcs[j + 1] + v.z*rcs[j + 2] + v.w*rcs[j + 3];
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0)
        {
            z[row] = acc;
            part = (double)acc*rc[row];
        }
    }
    __shared__ double red[CO_B/64];
    if (lane == 0) red[wid] = part;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < CO_B/64; w++) t += red[w];
    if (blockIdx.x == 0)
    {
        // the fine partials, fixed order
        double f = 0.0;
        for (int32_t i = threadIdx.x; i < nF; i += CO_B) f += fine[i];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) f += __shfl_xor(f, off, 64);
        __syncthreads();
        if (lane == 0) red[wid] = f;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 0; w < CO_B/64; w++) t += red[w];
    }
    partials_arrive(t, 0.0, partials, gridDim.x, blockIdx.x, F);
}

int32_t launch_coarse_apply(int32_t nT, int32_t ld, const float* inv, const double* rc,
                            float* z, const double* fine, int32_t nF, double* partials,
                            const FusedFinalize& F, const DevState* st, hipStream_t s)
{
    const int32_t nb = (nT + 3)/4;
    hipLaunchKernelGGL(k_coarse_apply, dim3(nb), dim3(CO_B), (size_t)ld*sizeof(float), s, nT,
                       ld, inv, rc, z, fine, nF, partials, F, st);
    return nb;
}

} // namespace fh
