// Fused vector kernels and the device-side scalar logic of the Krylov
// solvers.  Each vector kernel writes one deterministic partial sum per
// block; k_finalize adds the partials in a fixed order and applies the
// solver's scalar update to DevState, so a solve needs no host round trip
// per iteration.
#include "fh_kernels.h"
#include "fh_solver_kernels.h"
#include "fh_finalize_dev.h"
#include "fh_reduce_dev.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include <cfloat>
#include <cmath>

namespace fh
{

#define FH_BLOCK 256
// grid of the strided reduction kernels (few partials to sum)
#define MIX_BLOCKS 1024

// block sum of (a, b) over an FH_BLOCK block, then partials_arrive
__device__ __forceinline__ void write_partials_fin(double a, double b, double* partials,
                                                   const FusedFinalize& F)
{
    __shared__ double smem[2*FH_BLOCK/64];
    block_sum2(a, b, smem);
    partials_arrive(a, b, partials, gridDim.x, blockIdx.x, F);
}

// ---- reductions over vectors ----------------------------------------------

template<class Ta, class Tb>
__global__ void k_dot(int32_t n, const Ta* __restrict__ a, const Tb* __restrict__ b,
                      double* partials, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double v = 0.0;
    if (i < n) v = (double)a[i]*(double)b[i];
    write_partials(v, 0.0, partials, gridDim.x);
}

template<class T>
__global__ void k_sum_abs(int32_t n, const T* __restrict__ a, double* partials)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (i < n) { s = (double)a[i]; m = fabs((double)a[i]); }
    write_partials(s, m, partials, gridDim.x);
}

// lduMatrixSolver.C:174-197 and PCG.C:95-111: rA = source - Apsi,
// sum0 = sum |Apsi - xRef*sumA| + |source - xRef*sumA|, sum1 = sum |rA|
__global__ void k_init_residual(int32_t n, const double* __restrict__ Apsi,
                                const double* __restrict__ source,
                                const double* __restrict__ sumA, double* __restrict__ rA,
                                double* partials, const DevState* st)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double nf = 0.0, rm = 0.0;
    if (i < n)
    {
        const double t = sumA[i]*st->sum0;
        nf = fabs(Apsi[i] - t) + fabs(source[i] - t);
        const double r = source[i] - Apsi[i];
        rA[i] = r;
        rm = fabs(r);
    }
    write_partials(nf, rm, partials, gridDim.x);
}

// ---- PCG (PCG.C:129-189) -------------------------------------------------

// diagonal / none preconditioner fused with wArA = sum wA*rA
template<class T>
__global__ void k_precond_point(int32_t n, int32_t kind, const T* __restrict__ rD,
                                const T* __restrict__ r, T* __restrict__ w,
                                double* partials, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double v = 0.0;
    if (i < n)
    {
        const T wi = (kind == FH_PRECON_DIAGONAL) ? rD[i]*r[i] : r[i];
        w[i] = wi;
        v = (double)wi*(double)r[i];
    }
    write_partials(v, 0.0, partials, gridDim.x);
}

// pA = wA (first iteration) or wA + beta*pA
template<class T>
__global__ void k_pcg_p(int32_t n, const T* __restrict__ w, T* __restrict__ p,
                        const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (st->nIterations == 0) p[i] = w[i];
    else
    {
        const T beta = (T)st->beta;
        p[i] = w[i] + beta*p[i];
    }
}

// y = A x fused with sum0 = sum y*z, sum1 = sum y*v (v may be null; z null
// means z = y, i.e. sum y*y)
template<class T>
__global__ void k_spmv_dot(RowsView R, const T* __restrict__ diag, const T* __restrict__ val,
                           const T* __restrict__ x, T* __restrict__ y,
                           const T* z, const T* v,
                           double* partials, const DevState* st)
{
    if (st->done) return;
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    if (r < R.n)
    {
        const int32_t nEnt = R.rowInfo[r] >> 16;
        const int32_t base = sell_base2(R.sliceStart, r);
        T acc = diag[r]*x[r];
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            acc += val[e]*x[R.col[e]];
        }
        acc = add_interfaces(R, r, val, x, acc);
        y[r] = acc;
        s0 = (double)acc*(double)(z ? z[r] : acc);
        if (v) s1 = (double)acc*(double)v[r];
    }
    write_partials(s0, s1, partials, gridDim.x);
}

// The same product on a hex box from direction-packed coefficients (no
// row indirection): diag, then the lower neighbours k, j, i, then the upper
// neighbours i, j, k -- the row order of the general kernel, so the sums are
// the same.  When uI is null the matrix is symmetric and the upper
// coefficient of a face is read as the lower coefficient of its upper cell.
// Blocks are 64 (i) x 4 (j) cells of one k-plane; consecutive blocks of one
// XCD (hardware block id mod 8) walk a contiguous slab of the box, so the
// neighbouring planes a block reads are in that XCD's L2.
#define BOX_BX 64
#define BOX_BY 4

int32_t box_spmv_blocks(int32_t nx, int32_t ny, int32_t nz)
{
    const int64_t nb = (int64_t)((nx + BOX_BX - 1)/BOX_BX)*((ny + BOX_BY - 1)/BOX_BY)*nz;
    return (int32_t)(8*((nb + 7)/8));
}

__global__ __launch_bounds__(BOX_BX*BOX_BY)
void k_box_spmv_dot_f32(int32_t nx, int32_t ny, int32_t nz,
                        const float* __restrict__ diag,
                        const float* __restrict__ lK, const float* __restrict__ lJ,
                        const float* __restrict__ lI, const float* __restrict__ uI,
                        const float* __restrict__ uJ, const float* __restrict__ uK,
                        const float* __restrict__ x, float* __restrict__ y,
                        const float* z, double* partials, const DevState* st)
{
    if (st->done) return;
    const int32_t gx = (nx + BOX_BX - 1)/BOX_BX, gy = (ny + BOX_BY - 1)/BOX_BY;
    const int32_t per = gridDim.x/8;
    // XCD-aware swizzle: hardware block b runs on XCD b % 8
    const int32_t logical = (int32_t)(blockIdx.x % 8)*per + (int32_t)(blockIdx.x/8);
    const int32_t nReal = gx*gy*nz;
    double s0 = 0.0;
    if (logical < nReal)
    {
        const int32_t bx = logical % gx, by = (logical/gx) % gy, bz = logical/(gx*gy);
        const int32_t gi = bx*BOX_BX + (int32_t)(threadIdx.x % BOX_BX);
        const int32_t gj = by*BOX_BY + (int32_t)(threadIdx.x/BOX_BX);
        const int32_t gk = bz;
        if (gi < nx && gj < ny)
        {
            const int64_t sxy = (int64_t)nx*ny;
            const int64_t c = gi + (int64_t)nx*gj + sxy*gk;
            float acc = diag[c]*x[c];
            if (gk > 0) acc += lK[c]*x[c - sxy];
            if (gj > 0) acc += lJ[c]*x[c - nx];
            if (gi > 0) acc += lI[c]*x[c - 1];
            if (uI)
            {
                if (gi < nx - 1) acc += uI[c]*x[c + 1];
                if (gj < ny - 1) acc += uJ[c]*x[c + nx];
                if (gk < nz - 1) acc += uK[c]*x[c + sxy];
            }
            else
            {
                if (gi < nx - 1) acc += lI[c + 1]*x[c + 1];
                if (gj < ny - 1) acc += lJ[c + nx]*x[c + nx];
                if (gk < nz - 1) acc += lK[c + sxy]*x[c + sxy];
            }
            y[c] = acc;
            s0 = (double)acc*(double)(z ? z[c] : acc);
        }
    }
    write_partials(s0, 0.0, partials, gridDim.x);
}

// Four consecutive cells of one array: 16-byte (float) or two 16-byte
// (double) accesses; the index is a multiple of four
template<class T> struct Quad;
template<> struct Quad<float>
{
    __device__ static void load(const float* p, int64_t i, float (&o)[4])
    {
        const float4 q = *reinterpret_cast<const float4*>(p + i);
        o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w;
    }
    __device__ static void store(float* p, int64_t i, const float (&v)[4])
    {
        *reinterpret_cast<float4*>(p + i) = make_float4(v[0], v[1], v[2], v[3]);
    }
};
template<> struct Quad<double>
{
    __device__ static void load(const double* p, int64_t i, double (&o)[4])
    {
        const double2 a = *reinterpret_cast<const double2*>(p + i);
        const double2 b = *reinterpret_cast<const double2*>(p + i + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
    __device__ static void store(double* p, int64_t i, const double (&v)[4])
    {
        *reinterpret_cast<double2*>(p + i) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2*>(p + i + 2) = make_double2(v[2], v[3]);
    }
};

// Four cells along i per thread with vector loads (nx a multiple of 4):
// blocks of 64 threads x 4 rows = 256 x 4 cells of one k-plane, the same
// XCD-aware block order, and the same per-cell sums as above.
// FUSE_P: the product of p = w + beta*pOld (PCG.C:146-152, the pcg_p
// expression) computed on the fly from w and pOld for every cell the
// product reads; the new p of the block's cells is written to pNew (x is
// pOld here and z is pNew's values, i.e. the dot is p.Ap).
template<class T, bool FUSE_P>
__global__ __launch_bounds__(256)
void k_box_spmv_dot_x4(int32_t nx, int32_t ny, int32_t nz,
                       const T* __restrict__ diag,
                       const T* __restrict__ lK, const T* __restrict__ lJ,
                       const T* __restrict__ lI, const T* __restrict__ uI,
                       const T* __restrict__ uJ, const T* __restrict__ uK,
                       const T* __restrict__ x, T* __restrict__ y,
                       const T* z, double* partials, const DevState* st,
                       const T* __restrict__ wv, T* __restrict__ pNew)
{
    if (st->done) return;
    typedef Quad<T> Q;
    const bool first = FUSE_P && st->nIterations == 0;
    const T beta = FUSE_P ? (T)st->beta : T(0);
    const int32_t gx = (nx + 255)/256, gy = (ny + 3)/4;
    const int32_t nReal = gx*gy*nz;
    // XCD-aware order: the blocks of XCD (b mod 8) stride over one eighth
    // of the tiles, contiguous in k
    const int32_t slab = (nReal + 7)/8;
    const int32_t stride = gridDim.x/8;
    double s0 = 0.0;
    for (int32_t local = (int32_t)(blockIdx.x/8); local < slab; local += stride)
    {
        const int32_t logical = (int32_t)(blockIdx.x % 8)*slab + local;
        if (logical >= nReal) break;
        const int32_t bx = logical % gx, by = (logical/gx) % gy, bz = logical/(gx*gy);
        const int32_t i0 = bx*256 + 4*(int32_t)(threadIdx.x & 63);
        const int32_t gj = by*4 + (int32_t)(threadIdx.x >> 6);
        const int32_t gk = bz;
        if (i0 < nx && gj < ny)
        {
            const int64_t sxy = (int64_t)nx*ny;
            const int64_t c = i0 + (int64_t)nx*gj + sxy*gk;
            // the product's input at cells i..i+3 / cell i
            auto px4 = [&](int64_t i, T (&o)[4])
            {
                if (!FUSE_P) { Q::load(x, i, o); return; }
                Q::load(wv, i, o);
                if (first) return;
                T pq[4];
                Q::load(x, i, pq);
#pragma unroll
                for (int q = 0; q < 4; q++) o[q] = o[q] + beta*pq[q];
            };
            auto px1 = [&](int64_t i) -> T
            {
                if (!FUSE_P) return x[i];
                return first ? wv[i] : wv[i] + beta*x[i];
            };
            T X[4], dg[4], lk[4], lj[4], li[4], ui[4], uj[4], uk[4];
            T xk[4] = {}, xj[4] = {}, xjp[4] = {}, xkp[4] = {};
            px4(c, X);
            Q::load(diag, c, dg);
            Q::load(lK, c, lk); Q::load(lJ, c, lj); Q::load(lI, c, li);
            if (gk > 0) px4(c - sxy, xk);
            if (gj > 0) px4(c - nx, xj);
            if (gj < ny - 1) px4(c + nx, xjp);
            if (gk < nz - 1) px4(c + sxy, xkp);
            const T xm = i0 > 0 ? px1(c - 1) : T(0);
            const T xp = i0 + 4 < nx ? px1(c + 4) : T(0);
            if (FUSE_P) Q::store(pNew, c, X);
            if (uI)
            {
                Q::load(uI, c, ui); Q::load(uJ, c, uj); Q::load(uK, c, uk);
            }
            else
            {
                // symmetric: the upper coefficient of a face is the lower
                // one of its upper cell
                ui[0] = li[1]; ui[1] = li[2]; ui[2] = li[3];
                ui[3] = i0 + 4 < nx ? lI[c + 4] : T(0);
                if (gj < ny - 1) Q::load(lJ, c + nx, uj);
                else { uj[0] = uj[1] = uj[2] = uj[3] = T(0); }
                if (gk < nz - 1) Q::load(lK, c + sxy, uk);
                else { uk[0] = uk[1] = uk[2] = uk[3] = T(0); }
            }
            const T xs[6] = {xm, X[0], X[1], X[2], X[3], xp};
            T out[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
            {
                const int32_t gi = i0 + q;
                T acc = dg[q]*xs[q + 1];
                if (gk > 0) acc += lk[q]*xk[q];
                if (gj > 0) acc += lj[q]*xj[q];
                if (gi > 0) acc += li[q]*xs[q];
                if (gi < nx - 1) acc += ui[q]*xs[q + 2];
                if (gj < ny - 1) acc += uj[q]*xjp[q];
                if (gk < nz - 1) acc += uk[q]*xkp[q];
                out[q] = acc;
            }
            Q::store(y, c, out);
            T Z[4];
            if (FUSE_P) { for (int q = 0; q < 4; q++) Z[q] = X[q]; }
            else if (z) Q::load(z, c, Z);
            else { for (int q = 0; q < 4; q++) Z[q] = out[q]; }
            s0 += ((double)out[0]*(double)Z[0] + (double)out[1]*(double)Z[1])
               + ((double)out[2]*(double)Z[2] + (double)out[3]*(double)Z[3]);
        }
    }
    write_partials(s0, 0.0, partials, gridDim.x);
}

// The same fused product marching along k: a thread owns four cells along i
// of one row (i0..i0+3, j) for a run of kc planes and keeps the product's
// input of planes k-1, k, k+1 in registers (and, symmetric, the lower k
// coefficient of plane k+1, which is the upper one of plane k), so every
// value is read once from memory; only the j-1 / j+1 rows and the two i
// neighbours come from the caches.  Blocks are 64 x 4 rows; an XCD
// (hardware block id mod 8) takes a contiguous range of (row block, plane
// run) pairs, so the rows a block shares with its neighbours sit in that
// XCD's L2.  Same per-cell sums as k_box_spmv_dot_x4.
template<class T, bool FUSE_P>
__global__ __launch_bounds__(256)
void k_box_spmv_march(int32_t nx, int32_t ny, int32_t nz, int32_t kc,
                      const T* __restrict__ diag,
                      const T* __restrict__ lK, const T* __restrict__ lJ,
                      const T* __restrict__ lI, const T* __restrict__ uI,
                      const T* __restrict__ uJ, const T* __restrict__ uK,
                      const T* __restrict__ x, T* __restrict__ y,
                      const T* z, double* partials, const DevState* st,
                      const T* __restrict__ wv, T* __restrict__ pNew)
{
    if (st->done) return;
    typedef Quad<T> Q;
    const bool first = FUSE_P && st->nIterations == 0;
    const T beta = FUSE_P ? (T)st->beta : T(0);
    const int32_t nIB = (nx + 255)/256, nJB = (ny + 3)/4, nKB = (nz + kc - 1)/kc;
    const int32_t nReal = nIB*nJB*nKB;
    const int32_t slab = (nReal + 7)/8;
    const int32_t stride = gridDim.x/8;
    const int64_t sxy = (int64_t)nx*ny;
    double s0 = 0.0;
    for (int32_t local = (int32_t)(blockIdx.x/8); local < slab; local += stride)
    {
        const int32_t logical = (int32_t)(blockIdx.x % 8)*slab + local;
        if (logical >= nReal) break;
        const int32_t ib = logical % nIB, jb = (logical/nIB) % nJB, kb = logical/(nIB*nJB);
        const int32_t i0 = ib*256 + 4*(int32_t)(threadIdx.x & 63);
        const int32_t gj = jb*4 + (int32_t)(threadIdx.x >> 6);
        if (i0 >= nx || gj >= ny) continue;
        const int32_t k0 = kb*kc, k1 = min(nz, k0 + kc);
        const int64_t rowc = i0 + (int64_t)nx*gj;
        // the product's input at cells i..i+3 / cell i
        auto px4 = [&](int64_t i, T (&o)[4])
        {
            if (!FUSE_P) { Q::load(x, i, o); return; }
            Q::load(wv, i, o);
            if (first) return;
            T pq[4];
            Q::load(x, i, pq);
#pragma unroll
            for (int q = 0; q < 4; q++) o[q] = o[q] + beta*pq[q];
        };
        auto px1 = [&](int64_t i) -> T
        {
            if (!FUSE_P) return x[i];
            return first ? wv[i] : wv[i] + beta*x[i];
        };
        T xm[4] = {}, x0[4], xn[4] = {}, lkc[4], lkn[4] = {};
        if (k0 > 0) px4(rowc + sxy*(k0 - 1), xm);
        px4(rowc + sxy*k0, x0);
        Q::load(lK, rowc + sxy*k0, lkc);
        for (int32_t k = k0; k < k1; k++)
        {
            const int64_t c = rowc + sxy*k;
            const bool up = k < nz - 1;
            if (up)
            {
                px4(c + sxy, xn);
                if (!uI) Q::load(lK, c + sxy, lkn);
            }
            T dg[4], lj[4], li[4], ui[4], uj[4] = {}, uk[4] = {};
            T xj[4] = {}, xjp[4] = {};
            Q::load(diag, c, dg);
            Q::load(lJ, c, lj);
            Q::load(lI, c, li);
            if (gj > 0) px4(c - nx, xj);
            if (gj < ny - 1) px4(c + nx, xjp);
            const T xw = i0 > 0 ? px1(c - 1) : T(0);
            const T xe = i0 + 4 < nx ? px1(c + 4) : T(0);
            if (uI)
            {
                Q::load(lK, c, lkc);
                Q::load(uI, c, ui); Q::load(uJ, c, uj); Q::load(uK, c, uk);
            }
            else
            {
                // symmetric: the upper coefficient of a face is the lower
                // one of its upper cell
                ui[0] = li[1]; ui[1] = li[2]; ui[2] = li[3];
                ui[3] = i0 + 4 < nx ? lI[c + 4] : T(0);
                if (gj < ny - 1) Q::load(lJ, c + nx, uj);
#pragma unroll
                for (int q = 0; q < 4; q++) uk[q] = lkn[q];
            }
            if (FUSE_P) Q::store(pNew, c, x0);
            const T xs[6] = {xw, x0[0], x0[1], x0[2], x0[3], xe};
            T out[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
            {
                const int32_t gi = i0 + q;
                T acc = dg[q]*xs[q + 1];
                if (k > 0) acc += lkc[q]*xm[q];
                if (gj > 0) acc += lj[q]*xj[q];
                if (gi > 0) acc += li[q]*xs[q];
                if (gi < nx - 1) acc += ui[q]*xs[q + 2];
                if (gj < ny - 1) acc += uj[q]*xjp[q];
                if (up) acc += uk[q]*xn[q];
                out[q] = acc;
            }
            Q::store(y, c, out);
            T Z[4];
            if (FUSE_P) { for (int q = 0; q < 4; q++) Z[q] = x0[q]; }
            else if (z) Q::load(z, c, Z);
            else { for (int q = 0; q < 4; q++) Z[q] = out[q]; }
            s0 += ((double)out[0]*(double)Z[0] + (double)out[1]*(double)Z[1])
               + ((double)out[2]*(double)Z[2] + (double)out[3]*(double)Z[3]);
#pragma unroll
            for (int q = 0; q < 4; q++)
            {
                xm[q] = x0[q];
                x0[q] = xn[q];
                lkc[q] = lkn[q];
                xn[q] = T(0);
                lkn[q] = T(0);
            }
        }
    }
    write_partials(s0, 0.0, partials, gridDim.x);
}

// The fused PCG product over slabs of ROWS rows: a block of 64 x ROWS
// threads owns a 256 (i) x ROWS (j) tile of every plane of a run of kc
// planes and marches along k as k_box_spmv_march does (planes k-1, k, k+1 of
// the product's input and the lower-k coefficient of plane k+1 in
// registers), but the j-1 / j+1 rows and the j+1 row of the lower-j
// coefficients come from the block's own LDS copy of the plane; only the two
// rows just outside the slab are loaded from memory.  An XCD (block id mod
// 8) takes a contiguous range of slabs, so those two rows are usually in its
// L2.  Same per-cell sums as k_box_spmv_dot_x4.
template<class T, int ROWS>
__global__ __launch_bounds__(64*ROWS)
void k_box_spmv_slab(int32_t nx, int32_t ny, int32_t nz, int32_t kc,
                     const T* __restrict__ diag,
                     const T* __restrict__ lK, const T* __restrict__ lJ,
                     const T* __restrict__ lI, const T* __restrict__ uI,
                     const T* __restrict__ uJ, const T* __restrict__ uK,
                     const T* __restrict__ pOld, T* __restrict__ y,
                     double* partials, const DevState* st,
                     const T* __restrict__ wv, T* __restrict__ pNew, FusedFinalize F,
                     const T* __restrict__ zc, int32_t nTI, int32_t nTJ)
{
    if (st->done) return;
    typedef Quad<T> Q;
    // plane buffers: rows -1 .. ROWS of the product's input and, symmetric,
    // of the lower-j coefficients (row r at index r + 1)
    __shared__ T sX[ROWS + 2][256];
    __shared__ T sJ[ROWS + 2][256];
    __shared__ double red[ROWS];
    const bool first = st->nIterations == 0;
    const T beta = (T)st->beta;
    const int32_t nIB = (nx + 255)/256, nJB = (ny + ROWS - 1)/ROWS, nKB = (nz + kc - 1)/kc;
    const int32_t nReal = nIB*nJB*nKB;
    const int32_t slab = (nReal + 7)/8;
    const int32_t stride = gridDim.x/8;
    const int64_t sxy = (int64_t)nx*ny;
    const int q = threadIdx.x & 63, r = threadIdx.x >> 6;
    const bool sym = uI == nullptr;
    double s0 = 0.0;
    for (int32_t local = (int32_t)(blockIdx.x/8); local < slab; local += stride)
    {
        const int32_t logical = (int32_t)(blockIdx.x % 8)*slab + local;
        if (logical >= nReal) break;        // uniform over the block
        const int32_t ib = logical % nIB, jb = (logical/nIB) % nJB, kb = logical/(nIB*nJB);
        const int32_t i0 = ib*256 + 4*q;
        const int32_t gj = jb*ROWS + r;
        const bool ok = i0 < nx && gj < ny;
        const int32_t k0 = kb*kc, k1 = min(nz, k0 + kc);
        // rows just outside the slab, loaded by the first / last wave
        const int32_t hj = r == 0 ? jb*ROWS - 1 : jb*ROWS + ROWS;
        const bool halo = (r == 0 || r == ROWS - 1) && i0 < nx && hj >= 0 && hj < ny;
        const int hslot = r == 0 ? 0 : ROWS + 1;
        const int64_t rowc = (int64_t)(ok ? i0 : 0) + (int64_t)nx*(ok ? gj : 0);
        const int64_t rowh = (int64_t)(halo ? i0 : 0) + (int64_t)nx*(halo ? hj : 0);
        // zc: a coarse correction per 16^3 tile added to w (two-level
        // preconditioner): p = (w + zc) + beta*pOld
        auto zt = [&](int32_t xx, int32_t yy, int32_t kk) -> T
        {
            return zc ? zc[(xx >> 4) + nTI*((yy >> 4) + nTJ*(kk >> 4))] : T(0);
        };
        auto px4 = [&](int64_t i, int32_t yy, int32_t kk, T (&o)[4])
        {
            Q::load(wv, i, o);
            if (zc)
            {
                const T zv = zt(i0, yy, kk);
#pragma unroll
                for (int e = 0; e < 4; e++) o[e] = o[e] + zv;
            }
            if (first) return;
            T pq[4];
            Q::load(pOld, i, pq);
#pragma unroll
            for (int e = 0; e < 4; e++) o[e] = o[e] + beta*pq[e];
        };
        auto px1 = [&](int64_t i, int32_t xx, int32_t kk) -> T
        {
            const T wz = zc ? wv[i] + zt(xx, gj, kk) : wv[i];
            return first ? wz : wz + beta*pOld[i];
        };
        // every input of plane k+1 is loaded while plane k is computed
        T xm[4] = {}, x0[4] = {}, xn[4] = {}, lkc[4] = {}, lkn[4] = {};
        T dg[4] = {}, lj[4] = {}, li[4] = {}, hx[4] = {}, hl[4] = {};
        T dgn[4] = {}, ljn[4] = {}, lin[4] = {}, hxn[4] = {}, hln[4] = {};
        if (ok)
        {
            const int64_t c = rowc + sxy*k0;
            if (k0 > 0) px4(c - sxy, gj, k0 - 1, xm);
            px4(c, gj, k0, x0);
            Q::load(lK, c, lkc);
            Q::load(diag, c, dg);
            Q::load(lJ, c, lj);
            Q::load(lI, c, li);
        }
        if (halo)
        {
            px4(rowh + sxy*k0, hj, k0, hx);
            if (sym) Q::load(lJ, rowh + sxy*k0, hl);
        }
        for (int32_t k = k0; k < k1; k++)
        {
            const int64_t c = rowc + sxy*k;
            const bool up = k < nz - 1;
            const bool more = k + 1 < k1;
            if (ok && up)
            {
                px4(c + sxy, gj, k + 1, xn);
                if (sym) Q::load(lK, c + sxy, lkn);
                if (more)
                {
                    Q::load(diag, c + sxy, dgn);
                    Q::load(lJ, c + sxy, ljn);
                    Q::load(lI, c + sxy, lin);
                }
            }
            if (halo && more)
            {
                px4(rowh + sxy*(k + 1), hj, k + 1, hxn);
                if (sym) Q::load(lJ, rowh + sxy*(k + 1), hln);
            }
            // publish this plane's row (and the halo row) to the block
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                sX[r + 1][4*q + e] = x0[e];
                sJ[r + 1][4*q + e] = lj[e];
                if (halo) { sX[hslot][4*q + e] = hx[e]; sJ[hslot][4*q + e] = hl[e]; }
            }
            __syncthreads();
            if (ok)
            {
                T xj[4], xjp[4], uj[4] = {}, ui[4], uk[4], lkk[4];
#pragma unroll
                for (int e = 0; e < 4; e++)
                {
                    xj[e] = gj > 0 ? sX[r][4*q + e] : T(0);
                    xjp[e] = gj < ny - 1 ? sX[r + 2][4*q + e] : T(0);
                }
                // i neighbours: inside the 256-cell row from LDS, across
                // i-blocks from memory
                const T xw = i0 == 0 ? T(0)
                           : (q > 0 ? sX[r + 1][4*q - 1] : px1(c - 1, i0 - 1, k));
                const T xe = i0 + 4 >= nx ? T(0)
                           : (q < 63 ? sX[r + 1][4*q + 4] : px1(c + 4, i0 + 4, k));
                if (!sym)
                {
                    Q::load(lK, c, lkk);
                    Q::load(uI, c, ui); Q::load(uJ, c, uj); Q::load(uK, c, uk);
                }
                else
                {
#pragma unroll
                    for (int e = 0; e < 4; e++) lkk[e] = lkc[e];
                    // symmetric: the upper coefficient of a face is the
                    // lower one of its upper cell
                    ui[0] = li[1]; ui[1] = li[2]; ui[2] = li[3];
                    ui[3] = i0 + 4 < nx ? lI[c + 4] : T(0);
#pragma unroll
                    for (int e = 0; e < 4; e++)
                    {
                        uj[e] = gj < ny - 1 ? sJ[r + 2][4*q + e] : T(0);
                        uk[e] = lkn[e];
                    }
                }
                Q::store(pNew, c, x0);
                const T xs[6] = {xw, x0[0], x0[1], x0[2], x0[3], xe};
                T out[4];
#pragma unroll
                for (int e = 0; e < 4; e++)
                {
                    const int32_t gi = i0 + e;
                    T acc = dg[e]*xs[e + 1];
                    if (k > 0) acc += lkk[e]*xm[e];
                    if (gj > 0) acc += lj[e]*xj[e];
                    if (gi > 0) acc += li[e]*xs[e];
                    if (gi < nx - 1) acc += ui[e]*xs[e + 2];
                    if (gj < ny - 1) acc += uj[e]*xjp[e];
                    if (up) acc += uk[e]*xn[e];
                    out[e] = acc;
                }
                Q::store(y, c, out);
                s0 += ((double)out[0]*(double)x0[0] + (double)out[1]*(double)x0[1])
                   + ((double)out[2]*(double)x0[2] + (double)out[3]*(double)x0[3]);
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < 4; e++)
            {
                xm[e] = x0[e];
                x0[e] = xn[e];
                lkc[e] = lkn[e];
                dg[e] = dgn[e];
                lj[e] = ljn[e];
                li[e] = lin[e];
                hx[e] = hxn[e];
                hl[e] = hln[e];
                xn[e] = T(0);
                lkn[e] = T(0);
            }
        }
    }
    // one partial per block: wave butterflies, then the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s0 += __shfl_xor(s0, off, 64);
    if (q == 0) red[r] = s0;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < ROWS; w++) t += red[w];
    partials_arrive(t, 0.0, partials, gridDim.x, blockIdx.x, F);
}

// planes per marching run: about 1024 runs over the box, 2..32 planes each
static int32_t box_march_planes(int32_t nx, int32_t ny, int32_t nz)
{
    const int64_t rows = (int64_t)((nx + 255)/256)*((ny + 3)/4);
    const int64_t kc = (int64_t)nz*rows/1024;
    return (int32_t)std::max<int64_t>(2, std::min<int64_t>(32, kc));
}

static int32_t box_x4_blocks(int32_t nx, int32_t ny, int32_t nz)
{
    const int64_t nb4 = (int64_t)((nx + 255)/256)*((ny + 3)/4)*nz;
    return (int32_t)std::min<int64_t>(8*((nb4 + 7)/8), 2048);
}

template<class T>
int32_t launch_box_pcg_spmv(int32_t nx, int32_t ny, int32_t nz, const T* diag,
                            const T* const dir[6], bool symmetric, const T* w,
                            const T* pOld, T* pNew, T* y, double* partials,
                            const DevState* st, hipStream_t s, const FusedFinalize& F,
                            const T* zc)
{
    if (nx & 3) return -1;
    static const bool plane = [] {
        const char* e = std::getenv("FOAMHIP_BOX_SPMV");
        return e && std::strcmp(e, "plane") == 0;
    }();
    if (plane)
    {
        const int32_t nb = box_x4_blocks(nx, ny, nz);
        hipLaunchKernelGGL((k_box_spmv_dot_x4<T, true>), dim3(nb), dim3(256), 0, s, nx, ny, nz,
                           diag, dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                           symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], pOld, y,
                           nullptr, partials, st, w, pNew);
        return nb;
    }
    static const bool march = [] {
        const char* e = std::getenv("FOAMHIP_BOX_SPMV");
        return e && std::strcmp(e, "march") == 0;
    }();
    if (!march)
    {
        // slabs of 16 (FP32) / 8 (FP64) rows, about one block per CU
        constexpr int ROWS = sizeof(T) == 4 ? 16 : 8;
        const int64_t slabs = (int64_t)((nx + 255)/256)*((ny + ROWS - 1)/ROWS);
        const int32_t kc = (int32_t)std::max<int64_t>(
            4, std::min<int64_t>(64, (int64_t)nz*slabs/(sizeof(T) == 4 ? 256 : 512)));
        const int64_t runs = slabs*((nz + kc - 1)/kc);
        const int32_t nb = (int32_t)std::min<int64_t>(8*((runs + 7)/8), 2048);
        hipLaunchKernelGGL((k_box_spmv_slab<T, ROWS>), dim3(nb), dim3(64*ROWS), 0, s, nx, ny, nz,
                           kc, diag, dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                           symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], pOld, y,
                           partials, st, w, pNew, F, zc, (nx + 15)/16, (ny + 15)/16);
        return nb;
    }
    const int32_t kc = box_march_planes(nx, ny, nz);
    const int64_t runs = (int64_t)((nx + 255)/256)*((ny + 3)/4)*((nz + kc - 1)/kc);
    const int32_t nb = (int32_t)std::min<int64_t>(8*((runs + 7)/8), 2048);
    hipLaunchKernelGGL((k_box_spmv_march<T, true>), dim3(nb), dim3(256), 0, s, nx, ny, nz, kc,
                       diag, dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                       symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], pOld, y,
                       nullptr, partials, st, w, pNew);
    return nb;
}

template int32_t launch_box_pcg_spmv<float>(int32_t, int32_t, int32_t, const float*,
                                            const float* const[6], bool, const float*,
                                            const float*, float*, float*, double*,
                                            const DevState*, hipStream_t,
                                            const FusedFinalize&, const float*);
template int32_t launch_box_pcg_spmv<double>(int32_t, int32_t, int32_t, const double*,
                                             const double* const[6], bool, const double*,
                                             const double*, double*, double*, double*,
                                             const DevState*, hipStream_t,
                                             const FusedFinalize&, const double*);

int32_t launch_box_spmv_dot_f32(int32_t nx, int32_t ny, int32_t nz, const float* diag,
                                const float* const dir[6], bool symmetric, const float* x,
                                float* y, const float* z, double* partials,
                                const DevState* st, hipStream_t s)
{
    if ((nx & 3) == 0)
    {
        const int32_t nb = box_x4_blocks(nx, ny, nz);
        hipLaunchKernelGGL((k_box_spmv_dot_x4<float, false>), dim3(nb), dim3(256), 0, s, nx, ny, nz,
                           diag, dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                           symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], x, y, z,
                           partials, st, nullptr, nullptr);
        return nb;
    }
    const int32_t nb = box_spmv_blocks(nx, ny, nz);
    hipLaunchKernelGGL(k_box_spmv_dot_f32, dim3(nb), dim3(BOX_BX*BOX_BY), 0, s,
                       nx, ny, nz, diag, dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                       symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], x, y, z,
                       partials, st);
    return nb;
}

// psi += alpha*pA; rA -= alpha*wA; sum0 = sum |rA|
template<class T>
__global__ void k_pcg_update(int32_t n, T* __restrict__ psi, T* __restrict__ r,
                             const T* __restrict__ p, const T* __restrict__ w,
                             double* partials, const DevState* st, FusedFinalize F)
{
    if (st->done) return;
    const T alpha = (T)st->alpha;
    double s0 = 0.0;
    // one value per thread on a full grid; strided, four values per thread
    // in flight, on the short grid of the self-finalizing launch
    const int32_t stride = gridDim.x*FH_BLOCK;
    int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    for (; i + 3*stride < n; i += 4*stride)
    {
        T x4[4], r4[4], p4[4], w4[4];
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            x4[u] = psi[i + u*stride];
            r4[u] = r[i + u*stride];
            p4[u] = p[i + u*stride];
            w4[u] = w[i + u*stride];
        }
#pragma unroll
        for (int u = 0; u < 4; u++)
        {
            psi[i + u*stride] = x4[u] + alpha*p4[u];
            const T ri = r4[u] - alpha*w4[u];
            r[i + u*stride] = ri;
            s0 += fabs((double)ri);
        }
    }
    for (; i < n; i += stride)
    {
        psi[i] += alpha*p[i];
        const T ri = r[i] - alpha*w[i];
        r[i] = ri;
        s0 += fabs((double)ri);
    }
    write_partials_fin(s0, 0.0, partials, F);
}

// ---- PBiCGStab (PBiCGStab.C:149-248) --------------------------------------

// pA = rA (first) or rA + beta*(pA - omega*AyA)
template<class T>
__global__ void k_bicg_p(int32_t n, const T* __restrict__ r, T* __restrict__ p,
                         const T* __restrict__ AyA, const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i >= n) return;
    if (st->nIterations == 0) p[i] = r[i];
    else
    {
        const T beta = (T)st->bicgBeta, omega = (T)st->omega;
        p[i] = r[i] + beta*(p[i] - omega*AyA[i]);
    }
}

// sA = rA - alpha*AyA; sum0 = sum |sA|
template<class T>
__global__ void k_bicg_s(int32_t n, const T* __restrict__ r, const T* __restrict__ AyA,
                         T* __restrict__ s, double* partials, const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s0 = 0.0;
    if (i < n)
    {
        const T alpha = (T)st->bicgAlpha;
        const T si = r[i] - alpha*AyA[i];
        s[i] = si;
        s0 = fabs((double)si);
    }
    write_partials(s0, 0.0, partials, gridDim.x);
}

// early exit of PBiCGStab.C:204-219: psi += alpha*yA, applied only by the
// launch of the iteration that converged (seq match)
template<class T>
__global__ void k_bicg_early(int32_t n, T* __restrict__ psi, const T* __restrict__ y,
                             const DevState* st, int32_t seq)
{
    if (st->earlySeq != seq) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i >= n) return;
    const T alpha = (T)st->bicgAlpha;
    psi[i] += alpha*y[i];
}

// psi += alpha*yA + omega*zA; rA = sA - omega*tA; sum0 = |rA|, sum1 = rA0.rA
template<class T>
__global__ void k_bicg_update(int32_t n, T* __restrict__ psi, const T* __restrict__ y,
                              const T* __restrict__ z, T* __restrict__ r,
                              const T* __restrict__ s, const T* __restrict__ t,
                              const T* __restrict__ r0, double* partials, const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s0 = 0.0, s1 = 0.0;
    if (i < n)
    {
        const T alpha = (T)st->bicgAlpha, omega = (T)st->omega;
        psi[i] += alpha*y[i] + omega*z[i];
        const T ri = s[i] - omega*t[i];
        r[i] = ri;
        s0 = fabs((double)ri);
        s1 = (double)r0[i]*(double)ri;
    }
    write_partials(s0, s1, partials, gridDim.x);
}

// ---- finalize: fixed-order sum of the block partials, then scalar logic ---

// The partials are summed by 1024 threads, each with four independent
// accumulators over a fixed stride (loads in flight instead of one
// dependent chain), then across the block in a fixed order.
#define FIN_THREADS 1024
__global__ __launch_bounds__(FIN_THREADS)
void k_finalize(int32_t nBlocks, const double* partials, int32_t op,
                DevState* st, int32_t n, int32_t seq)
{
    if (op != FIN_SUM && op != FIN_XREF && op != FIN_INIT && op != FIN_INNER_INIT
        && op != FIN_OUTER_RES && st->done) return;
    // no residual replacement this iteration: nothing to sum
    if (op == FIN_MIXED_TRUE && !st->replace) return;
    __shared__ double smem[2*FIN_THREADS/64];
    const double* pa = partials;
    const double* pb = partials + nBlocks;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0, b3 = 0.0;
    int32_t i = threadIdx.x;
    for (; i + 3*FIN_THREADS < nBlocks; i += 4*FIN_THREADS)
    {
        a0 += pa[i]; a1 += pa[i + FIN_THREADS];
        a2 += pa[i + 2*FIN_THREADS]; a3 += pa[i + 3*FIN_THREADS];
        b0 += pb[i]; b1 += pb[i + FIN_THREADS];
        b2 += pb[i + 2*FIN_THREADS]; b3 += pb[i + 3*FIN_THREADS];
    }
    for (; i < nBlocks; i += FIN_THREADS)
    {
        a0 += pa[i];
        b0 += pb[i];
    }
    double a = (a0 + a1) + (a2 + a3);
    double b = (b0 + b1) + (b2 + b3);
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
        smem[FIN_THREADS/64 + wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int w = 0; w < FIN_THREADS/64; w++) { sa += smem[w]; sb += smem[FIN_THREADS/64 + w]; }
        a = sa;
        b = sb;
    }
    if (threadIdx.x != 0) return;
    finalize_apply(a, b, op, st, n, seq);
}


// a fixed grid of at most MIX_BLOCKS blocks strides over the vector, so that
// the finalize sums few partials (the order stays fixed)
// The solution update alpha*p accumulates in FP32 (xc) between residual
// replacements and is folded into the FP64 solution x on the iteration after
// each replacement (x += xc + alpha*p in FP64, xc = 0): the FP64 solution is
// read and written a few times per solve instead of every iteration.
// rPad (or null): r also written on rows padded to nxp cells (the input
// layout of the tile-local apply on boxes whose rows are not a multiple of
// 16 cells)
__device__ __forceinline__ void mix_pad_store(float* rPad, int32_t nx, int32_t nxp, int32_t i,
                                              float v)
{
    if (rPad) rPad[(i % nx) + (int64_t)(i/nx)*nxp] = v;
}

__global__ void k_mixed_update(int32_t n, double* __restrict__ x, float* __restrict__ xc,
                               float* __restrict__ r, const float* __restrict__ p,
                               const float* __restrict__ w, double* partials,
                               const DevState* st, FusedFinalize F, int32_t nx, int32_t nxp,
                               float* __restrict__ rPad)
{
    if (st->done) return;
    const double alpha = st->alpha;
    const float alphaF = (float)alpha;
    const bool fold = st->foldPending != 0;
    double s0 = 0.0;
    const int32_t stride = gridDim.x*FH_BLOCK;
    int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (fold)
    {
        for (; i < n; i += stride)
        {
            x[i] += (double)xc[i] + alpha*(double)p[i];
            xc[i] = 0.f;
            const float ri = r[i] - alphaF*w[i];
            r[i] = ri;
            mix_pad_store(rPad, nx, nxp, i, ri);
            s0 += fabs((double)ri);
        }
    }
    else
    {
        // four values per thread in flight
        for (; i + 3*stride < n; i += 4*stride)
        {
            float c[4], rr[4], pp[4], ww[4];
#pragma unroll
            for (int u = 0; u < 4; u++)
            {
                c[u] = xc[i + u*stride];
                rr[u] = r[i + u*stride];
                pp[u] = p[i + u*stride];
                ww[u] = w[i + u*stride];
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
            {
                xc[i + u*stride] = c[u] + alphaF*pp[u];
                const float ri = rr[u] - alphaF*ww[u];
                r[i + u*stride] = ri;
                mix_pad_store(rPad, nx, nxp, i + u*stride, ri);
                s0 += fabs((double)ri);
            }
        }
        for (; i < n; i += stride)
        {
            xc[i] += alphaF*p[i];
            const float ri = r[i] - alphaF*w[i];
            r[i] = ri;
            mix_pad_store(rPad, nx, nxp, i, ri);
            s0 += fabs((double)ri);
        }
    }
    write_partials_fin(s0, 0.0, partials, F);
}

__global__ void k_mixed_replace(RowsView R, const double* __restrict__ diag,
                                const double* __restrict__ val, const double* __restrict__ x,
                                const float* __restrict__ xc,
                                const double* __restrict__ b, float* __restrict__ r32,
                                double* partials, const DevState* st, FusedFinalize F)
{
    // launched every iteration, works only on replacement iterations: a
    // strided grid of at most MIX_BLOCKS blocks keeps the idle launch short
    if (st->done || !st->replace) return;
    double s1 = 0.0;
    for (int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x; r < R.n; r += gridDim.x*FH_BLOCK)
    {
        const int32_t nEnt = R.rowInfo[r] >> 16;
        const int32_t base = sell_base2(R.sliceStart, r);
        // the current solution: x plus the FP32 correction not yet folded
        double acc = b[r] - diag[r]*(x[r] + (double)xc[r]);
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            const int32_t c = R.col[e];
            acc -= val[e]*(x[c] + (double)xc[c]);
        }
        if (R.ifStart)
            for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++)
            {
                const int32_t c = R.ifNbr[e];
                acc -= val[R.ifBase + e]*(x[c] + (double)xc[c]);
            }
        r32[r] = (float)acc;
        s1 += fabs(acc);
    }
    write_partials_fin(0.0, s1, partials, F);
}

// k_mixed_replace on a box from the packed FP64 coefficients (lK, lJ, lI
// and, asymmetric, uI, uJ, uK; symmetric: the upper coefficient of a cell is
// the lower one of its upper neighbour).  The row order of the SELL rows
// (diag, -k, -j, -i, +i, +j, +k), the same cells per thread and the same
// partials: the result and its sum are that kernel's to the bit, without
// the column indices and the padded SELL values.
__global__ void k_mixed_replace_box(int32_t nx, int32_t ny, int32_t nz,
                                    const double* __restrict__ diag,
                                    const double* __restrict__ lK, const double* __restrict__ lJ,
                                    const double* __restrict__ lI, const double* __restrict__ uI,
                                    const double* __restrict__ uJ, const double* __restrict__ uK,
                                    const double* __restrict__ x, const float* __restrict__ xc,
                                    const double* __restrict__ b, float* __restrict__ r32,
                                    double* partials, const DevState* st, FusedFinalize F,
                                    int32_t nxp, float* __restrict__ rPad)
{
    if (st->done || !st->replace) return;
    const int32_t sy = nx, sz = nx*ny, n = sz*nz;
    double s1 = 0.0;
    for (int32_t c = blockIdx.x*FH_BLOCK + threadIdx.x; c < n; c += gridDim.x*FH_BLOCK)
    {
        const int32_t q = c/nx, gk = q/ny, gi = c - q*nx, gj = q - gk*ny;
        double acc = b[c] - diag[c]*(x[c] + (double)xc[c]);
        if (gk > 0) acc -= lK[c]*(x[c - sz] + (double)xc[c - sz]);
        if (gj > 0) acc -= lJ[c]*(x[c - sy] + (double)xc[c - sy]);
        if (gi > 0) acc -= lI[c]*(x[c - 1] + (double)xc[c - 1]);
        if (gi < nx - 1) acc -= (uI ? uI[c] : lI[c + 1])*(x[c + 1] + (double)xc[c + 1]);
        if (gj < ny - 1) acc -= (uJ ? uJ[c] : lJ[c + sy])*(x[c + sy] + (double)xc[c + sy]);
        if (gk < nz - 1) acc -= (uK ? uK[c] : lK[c + sz])*(x[c + sz] + (double)xc[c + sz]);
        r32[c] = (float)acc;
        mix_pad_store(rPad, nx, nxp, c, (float)acc);
        s1 += fabs(acc);
    }
    write_partials_fin(0.0, s1, partials, F);
}

int32_t launch_mixed_replace_box(int32_t nx, int32_t ny, int32_t nz, const double* diag,
                                 const double* const dir[6], bool symmetric, const double* x,
                                 const float* xc, const double* b, float* r32,
                                 double* partials, const DevState* st, hipStream_t s,
                                 const FusedFinalize& F, int32_t nxp, float* rPad)
{
    const int32_t nb = std::min(reduction_blocks(nx*ny*nz), MIX_BLOCKS);
    hipLaunchKernelGGL(k_mixed_replace_box, dim3(nb), dim3(FH_BLOCK), 0, s, nx, ny, nz, diag,
                       dir[0], dir[1], dir[2], symmetric ? nullptr : dir[3],
                       symmetric ? nullptr : dir[4], symmetric ? nullptr : dir[5], x, xc, b,
                       r32, partials, st, F, nxp, rPad);
    return nb;
}

__global__ void k_mixed_setup(DevState* st, double ratio)
{
    st->resLast = st->initialResidual;
    st->replaceRatio = ratio;
    st->replace = 0;
    st->nReplace = 0;
    st->foldPending = 0;
}

void launch_mixed_setup(DevState* st, double replaceRatio, hipStream_t s)
{
    hipLaunchKernelGGL(k_mixed_setup, dim3(1), dim3(1), 0, s, st, replaceRatio);
}

int32_t launch_mixed_update(int32_t n, double* x, float* xc, float* r, const float* p,
                            const float* w, double* partials, const DevState* st,
                            hipStream_t s, const FusedFinalize& F, int32_t nx, int32_t nxp,
                            float* rPad)
{
    const int32_t nb = std::min(reduction_blocks(n), MIX_BLOCKS);
    hipLaunchKernelGGL(k_mixed_update, dim3(nb), dim3(FH_BLOCK), 0, s,
                       n, x, xc, r, p, w, partials, st, F, nx, nxp, rPad);
    return nb;
}

int32_t launch_mixed_replace(const RowsView& R, const double* diag, const double* val,
                             const double* x, const float* xc, const double* b, float* r32,
                             double* partials, const DevState* st, hipStream_t s,
                             const FusedFinalize& F)
{
    const int32_t nb = std::min(reduction_blocks(R.n), MIX_BLOCKS);
    hipLaunchKernelGGL(k_mixed_replace, dim3(nb), dim3(FH_BLOCK), 0, s,
                       R, diag, val, x, xc, b, r32, partials, st, F);
    return nb;
}

__global__ void k_axpy_inplace(int32_t n, double* __restrict__ psi,
                               const double* __restrict__ w, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) psi[i] += w[i];
}

template<class T>
__global__ void k_sum_abs_state(int32_t n, const T* __restrict__ a, double* partials,
                                const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s = 0.0, m = 0.0;
    if (i < n) { s = (double)a[i]; m = fabs((double)a[i]); }
    write_partials(s, m, partials, gridDim.x);
}

template<class T>
__global__ void k_axpy_alpha(int32_t n, T* __restrict__ y, const T* __restrict__ x,
                             const DevState* st)
{
    if (st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) y[i] -= (T)st->alpha*x[i];
}

__global__ void k_sub_sum_abs(int32_t n, const double* __restrict__ Apsi,
                              const double* __restrict__ source, double* __restrict__ r,
                              double* partials, const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double m = 0.0;
    if (i < n)
    {
        const double v = source[i] - Apsi[i];
        r[i] = v;
        m = fabs(v);
    }
    write_partials(0.0, m, partials, gridDim.x);
}

__global__ void k_scale_convert(int32_t n, const double* __restrict__ x, double scale,
                                float* __restrict__ y)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) y[i] = (float)(x[i]*scale);
}

__global__ void k_axpy_f32(int32_t n, double* __restrict__ psi, const float* __restrict__ d,
                           double scale)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) psi[i] += scale*(double)d[i];
}

// diagonalSolver.C:70: psi = source/diag
__global__ void k_divide(int32_t n, const double* __restrict__ src,
                         const double* __restrict__ diag, double* __restrict__ psi)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) psi[i] = src[i]/diag[i];
}

// ---- launchers ---------------------------------------------------------------

void launch_axpy_inplace(int32_t n, double* psi, const double* w, const DevState* st,
                         hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_inplace, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, psi, w, st);
}

template<class T>
void launch_sum_abs_state(int32_t n, const T* a, double* partials, const DevState* st,
                          hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_abs_state<T>, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, a, partials, st);
}
template void launch_sum_abs_state<double>(int32_t, const double*, double*, const DevState*, hipStream_t);
template void launch_sum_abs_state<float>(int32_t, const float*, double*, const DevState*, hipStream_t);

void launch_finalize_smooth(int32_t n, const double* partials, DevState* st,
                            int32_t nSweeps, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s,
                       reduction_blocks(n), partials, (int32_t)FIN_SMOOTH_RES, st, n, nSweeps);
}

template<class T>
void launch_axpy_alpha(int32_t n, T* y, const T* x, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_alpha<T>, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, y, x, st);
}
template void launch_axpy_alpha<double>(int32_t, double*, const double*, const DevState*, hipStream_t);
template void launch_axpy_alpha<float>(int32_t, float*, const float*, const DevState*, hipStream_t);

void launch_sub_sum_abs(int32_t n, const double* Apsi, const double* source, double* r,
                        double* partials, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_sub_sum_abs, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, Apsi, source, r, partials, st);
}

void launch_scale_convert(int32_t n, const double* x, double scale, float* y, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_convert, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, x, scale, y);
}

void launch_axpy_f32(int32_t n, double* psi, const float* d, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_axpy_f32, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, psi, d, scale);
}

void launch_divide(int32_t n, const double* src, const double* diag, double* psi,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_divide, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s,
                       n, src, diag, psi);
}

static inline dim3 grid_for(int32_t n) { return dim3(reduction_blocks(n)); }

void launch_finalize(int32_t n, const double* partials, int32_t op, DevState* st,
                     int32_t seq, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s,
                       reduction_blocks(n), partials, op, st, n, seq);
}

void launch_finalize_blocks(int32_t nBlocks, int32_t n, const double* partials, int32_t op,
                            DevState* st, int32_t seq, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(FIN_THREADS), 0, s, nBlocks, partials, op, st, n,
                       seq);
}

template<class Ta, class Tb>
void launch_dot(int32_t n, const Ta* a, const Tb* b, double* partials, const DevState* st,
                hipStream_t s)
{
    hipLaunchKernelGGL((k_dot<Ta, Tb>), grid_for(n), dim3(FH_BLOCK), 0, s, n, a, b, partials, st);
}

template<class T>
void launch_sum_abs(int32_t n, const T* a, double* partials, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_abs<T>, grid_for(n), dim3(FH_BLOCK), 0, s, n, a, partials);
}

void launch_init_residual(int32_t n, const double* Apsi, const double* source,
                          const double* sumA, double* rA, double* partials,
                          const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_init_residual, grid_for(n), dim3(FH_BLOCK), 0, s,
                       n, Apsi, source, sumA, rA, partials, st);
}

template<class T>
void launch_precond_point(int32_t n, int32_t kind, const T* rD, const T* r, T* w,
                          double* partials, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_precond_point<T>, grid_for(n), dim3(FH_BLOCK), 0, s,
                       n, kind, rD, r, w, partials, st);
}

template<class T>
void launch_pcg_p(int32_t n, const T* w, T* p, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_p<T>, grid_for(n), dim3(FH_BLOCK), 0, s, n, w, p, st);
}

template<class T>
void launch_spmv_dot(const RowsView& R, const T* diag, const T* val, const T* x, T* y,
                     const T* z, const T* v, double* partials, const DevState* st,
                     hipStream_t s)
{
    hipLaunchKernelGGL(k_spmv_dot<T>, grid_for(R.n), dim3(FH_BLOCK), 0, s,
                       R, diag, val, x, y, z, v, partials, st);
}

template<class T>
void launch_pcg_update(int32_t n, T* psi, T* r, const T* p, const T* w, double* partials,
                       const DevState* st, hipStream_t s, const FusedFinalize& F)
{
    const dim3 grid = F.counter ? dim3(std::min(reduction_blocks(n), MIX_BLOCKS)) : grid_for(n);
    hipLaunchKernelGGL(k_pcg_update<T>, grid, dim3(FH_BLOCK), 0, s,
                       n, psi, r, p, w, partials, st, F);
}

// the same update without a solver state (measurement)
template<class T>
__global__ void k_pcg_update_nostate(int32_t n, T* __restrict__ psi, T* __restrict__ r,
                                     const T* __restrict__ p, const T* __restrict__ w,
                                     double* partials)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    double s0 = 0.0;
    if (i < n)
    {
        const T alpha = T(1e-3);
        psi[i] += alpha*p[i];
        const T ri = r[i] - alpha*w[i];
        r[i] = ri;
        s0 = fabs((double)ri);
    }
    write_partials(s0, 0.0, partials, gridDim.x);
}

template<class T>
void launch_pcg_update_nostate(int32_t n, T* psi, T* r, const T* p, const T* w,
                               double* partials, hipStream_t s)
{
    hipLaunchKernelGGL(k_pcg_update_nostate<T>, grid_for(n), dim3(FH_BLOCK), 0, s,
                       n, psi, r, p, w, partials);
}
template void launch_pcg_update_nostate<double>(int32_t, double*, double*, const double*, const double*, double*, hipStream_t);
template void launch_pcg_update_nostate<float>(int32_t, float*, float*, const float*, const float*, double*, hipStream_t);

template<class T>
void launch_bicg_p(int32_t n, const T* r, T* p, const T* AyA, const DevState* st,
                   hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_p<T>, grid_for(n), dim3(FH_BLOCK), 0, s, n, r, p, AyA, st);
}

template<class T>
void launch_bicg_s(int32_t n, const T* r, const T* AyA, T* sA, double* partials,
                   const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_s<T>, grid_for(n), dim3(FH_BLOCK), 0, s, n, r, AyA, sA,
                       partials, st);
}

template<class T>
void launch_bicg_early(int32_t n, T* psi, const T* y, const DevState* st, int32_t seq,
                       hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_early<T>, grid_for(n), dim3(FH_BLOCK), 0, s, n, psi, y, st, seq);
}

template<class T>
void launch_bicg_update(int32_t n, T* psi, const T* y, const T* z, T* r, const T* sA,
                        const T* tA, const T* r0, double* partials, const DevState* st,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_bicg_update<T>, grid_for(n), dim3(FH_BLOCK), 0, s,
                       n, psi, y, z, r, sA, tA, r0, partials, st);
}

#define FH_INST(T)                                                                   \
    template void launch_dot<T, T>(int32_t, const T*, const T*, double*, const DevState*, hipStream_t); \
    template void launch_sum_abs<T>(int32_t, const T*, double*, hipStream_t);        \
    template void launch_precond_point<T>(int32_t, int32_t, const T*, const T*, T*, double*, const DevState*, hipStream_t); \
    template void launch_pcg_p<T>(int32_t, const T*, T*, const DevState*, hipStream_t); \
    template void launch_spmv_dot<T>(const RowsView&, const T*, const T*, const T*, T*, const T*, const T*, double*, const DevState*, hipStream_t); \
    template void launch_pcg_update<T>(int32_t, T*, T*, const T*, const T*, double*, const DevState*, hipStream_t, const FusedFinalize&); \
    template void launch_bicg_p<T>(int32_t, const T*, T*, const T*, const DevState*, hipStream_t); \
    template void launch_bicg_s<T>(int32_t, const T*, const T*, T*, double*, const DevState*, hipStream_t); \
    template void launch_bicg_early<T>(int32_t, T*, const T*, const DevState*, int32_t, hipStream_t); \
    template void launch_bicg_update<T>(int32_t, T*, const T*, const T*, T*, const T*, const T*, const T*, double*, const DevState*, hipStream_t);

FH_INST(double)
FH_INST(float)

} // namespace fh
