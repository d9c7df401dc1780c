// Tile-local (block-Jacobi over 16 x 16 x bz tiles) DIC / DILU application
// in FP32 for the inner iterations of the mixed-precision solvers on hex
// boxes.  Arithmetic as in the tiled sweep's LOCAL mode; the data path is
// built for bandwidth: the coefficients of each direction are packed per
// cell (no row indirection) and every lane reads and writes whole i-rows of
// 16 contiguous cells.
#include "fh_tile_sweep.h"
#include "fh_finalize_dev.h"

#include <hip/hip_fp16.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>


namespace fh
{

#define TL_B 256

// Per-cell coefficient gather: for each cell of the box, the coefficient of
// its lower (K, J, I) and upper (I, J, K) neighbour entries from the SELL
// values; 0 where the neighbour does not exist.  With three destinations only
// the lower ones (a symmetric matrix: the upper coefficient of a cell is the
// lower one of its upper neighbour).
template<class T>
__global__ void k_pack_dir_coefs(int32_t nx, int32_t ny, int32_t nz, RowsView R,
                                 const T* __restrict__ val, T* __restrict__ lK,
                                 T* __restrict__ lJ, T* __restrict__ lI,
                                 T* __restrict__ uI, T* __restrict__ uJ,
                                 T* __restrict__ uK)
{
    const int64_t c = (int64_t)blockIdx.x*TL_B + threadIdx.x;
    const int64_t n = (int64_t)nx*ny*nz;
    if (c >= n) return;
    const int32_t gi = (int32_t)(c % nx), gj = (int32_t)((c/nx) % ny), gk = (int32_t)(c/((int64_t)nx*ny));
    const int32_t base = R.sliceStart[c >> 6] + (int32_t)(c & 63);
    int e = 0;
    T vK = 0, vJ = 0, vI = 0;
    if (gk > 0) vK = val[base + ((e++) << 6)];
    if (gj > 0) vJ = val[base + ((e++) << 6)];
    if (gi > 0) vI = val[base + ((e++) << 6)];
    lK[c] = vK; lJ[c] = vJ; lI[c] = vI;
    if (!uI) return;
    T wI = 0, wJ = 0, wK = 0;
    if (gi < nx - 1) wI = val[base + ((e++) << 6)];
    if (gj < ny - 1) wJ = val[base + ((e++) << 6)];
    if (gk < nz - 1) wK = val[base + ((e++) << 6)];
    uI[c] = wI; uJ[c] = wJ; uK[c] = wK;
}

template<class T>
void launch_pack_dir_coefs(const TileView& G, const RowsView& R, const T* val, int nDst,
                           T* const dst[6], hipStream_t s)
{
    const int64_t n = (int64_t)G.nx*G.ny*G.nz;
    T* up[3] = {nullptr, nullptr, nullptr};
    if (nDst == 6) { up[0] = dst[3]; up[1] = dst[4]; up[2] = dst[5]; }
    else if (nDst != 3) fail(FH_ERR_INVALID, "internal: %d direction arrays", nDst);
    hipLaunchKernelGGL(k_pack_dir_coefs<T>, dim3((unsigned)((n + TL_B - 1)/TL_B)), dim3(TL_B), 0,
                       s, G.nx, G.ny, G.nz, R, val, dst[0], dst[1], dst[2], up[0], up[1], up[2]);
}

template void launch_pack_dir_coefs<float>(const TileView&, const RowsView&, const float*, int,
                                           float* const[6], hipStream_t);
template void launch_pack_dir_coefs<double>(const TileView&, const RowsView&, const double*, int,
                                            double* const[6], hipStream_t);

// Row walk of a tile: a lane owns one i-row (j, k) of 16 cells and updates
// them in order along i (the i-neighbour in a register), the rows of one
// diagonal j + k run together, and a row reads the rows (j-1, k) and
// (j, k-1) of the previous diagonal from LDS.  Every row is read and written
// as 16 contiguous values.  One wavefront walks four tiles (16 lanes each),
// so its barrier is the wave's own.  Forward: w = rD*x - rD*(cK*w(k-1) +
// cJ*w(j-1) + cI*w(i-1)); backward (w in place): w -= rD*(cK*w(k+1) +
// cJ*w(j+1) + cI*w(i+1)) -- the reference's per-row order either way.
template<bool BWD, int BZ>
__global__ __launch_bounds__(64)
void k_tile_local_f32(TileView G, const float* a, const float* __restrict__ rD,
                      const float* __restrict__ cK, const float* __restrict__ cJ,
                      const float* __restrict__ cI, float* w, const DevState* st,
                      const float* __restrict__ r, double* partials)
{
    if (st && st->done) return;
    double dotWR = 0.0;      // backward with r: this lane's share of sum w*r
    __shared__ float buf[4][2][BZ][16];          // per tile: diagonal rows by k slot
    const int lane = threadIdx.x;
    const int g = lane >> 4, l16 = lane & 15;
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    const int tile = blockIdx.x*4 + g;
    const bool tileOk = tile < nTiles;
    const int t = tileOk ? tile : 0;
    const int TI = t % G.nTI, TJ = (t/G.nTI) % G.nTJ, TK = t/(G.nTI*G.nTJ);
    const int x0 = TI*16, y0 = TJ*16, z0 = TK*BZ;
    const int ex = min(16, G.nx - x0), ey = min(16, G.ny - y0), ez = min(BZ, G.nz - z0);
    const int64_t sy = G.nx, sz = (int64_t)G.nx*G.ny;
    const int64_t base = x0 + sy*y0 + sz*z0;
    const int nDiag = ey + ez - 1;
    // whole 16-wide rows starting on 16-byte boundaries: vector access
    const bool vec = (G.nx & 3) == 0 && ex == 16;
    // the wave walks the longest diagonal count of its tiles
    int nD = tileOk ? nDiag : 0;
    nD = max(nD, __shfl_xor(nD, 16, 64));
    nD = max(nD, __shfl_xor(nD, 32, 64));
    for (int step = 0; step < nD; step++)
    {
        const int d = BWD ? nDiag - 1 - step : step;
        // rows of diagonal d: k = kmin + l16, j = d - k
        const int kmin = max(0, d - (ey - 1));
        const int k = kmin + l16;
        const int j = d - k;
        const bool act = tileOk && step < nDiag && k < ez && j >= 0 && j < ey;
        float out[16];
        if (act)
        {
            const int64_t row = base + sy*j + sz*k;
            // the row's inputs: 16 contiguous values per array (4-wide
            // vector loads when every row starts 16-byte aligned)
            float A[16], D[16], CK[16], CJ[16], CI[16];
            if (vec)
            {
#pragma unroll
                for (int v = 0; v < 4; v++)
                {
                    const float4 qa = reinterpret_cast<const float4*>(a + row)[v];
                    const float4 qd = reinterpret_cast<const float4*>(rD + row)[v];
                    const float4 qk = reinterpret_cast<const float4*>(cK + row)[v];
                    const float4 qj = reinterpret_cast<const float4*>(cJ + row)[v];
                    const float4 qi = reinterpret_cast<const float4*>(cI + row)[v];
                    A[4*v] = qa.x; A[4*v + 1] = qa.y; A[4*v + 2] = qa.z; A[4*v + 3] = qa.w;
                    D[4*v] = qd.x; D[4*v + 1] = qd.y; D[4*v + 2] = qd.z; D[4*v + 3] = qd.w;
                    CK[4*v] = qk.x; CK[4*v + 1] = qk.y; CK[4*v + 2] = qk.z; CK[4*v + 3] = qk.w;
                    CJ[4*v] = qj.x; CJ[4*v + 1] = qj.y; CJ[4*v + 2] = qj.z; CJ[4*v + 3] = qj.w;
                    CI[4*v] = qi.x; CI[4*v + 1] = qi.y; CI[4*v + 2] = qi.z; CI[4*v + 3] = qi.w;
                }
            }
            else
            {
#pragma unroll
                for (int i = 0; i < 16; i++)
                {
                    const bool in = i < ex;
                    A[i] = in ? a[row + i] : 0.f;
                    D[i] = in ? rD[row + i] : 0.f;
                    CK[i] = in ? cK[row + i] : 0.f;
                    CJ[i] = in ? cJ[row + i] : 0.f;
                    CI[i] = in ? cI[row + i] : 0.f;
                }
            }
            const float* nb = buf[g][(step + 1) & 1][0];
            const bool hasK = BWD ? k < ez - 1 : k > 0;
            const bool hasJ = BWD ? j < ey - 1 : j > 0;
            // slots of the neighbour rows on the previous diagonal: (j -/+ 1, k)
            // keeps k, (j, k -/+ 1) moves one slot
            const int sK = BWD ? k + 1 : k - 1;
            float prev = 0.f;
#pragma unroll
            for (int q = 0; q < 16; q++)
            {
                const int i = BWD ? 15 - q : q;
                out[i] = 0.f;
                if (i < ex)
                {
                    const float Di = D[i];
                    float v = BWD ? A[i] : Di*A[i];
                    if (hasK) v -= Di*CK[i]*nb[sK*16 + i];
                    if (hasJ) v -= Di*CJ[i]*nb[k*16 + i];
                    if (BWD ? i < ex - 1 : i > 0) v -= Di*CI[i]*prev;
                    prev = v;
                    out[i] = v;
                }
            }
            if (vec)
            {
#pragma unroll
                for (int v = 0; v < 4; v++)
                    reinterpret_cast<float4*>(w + row)[v] =
                        make_float4(out[4*v], out[4*v + 1], out[4*v + 2], out[4*v + 3]);
            }
            else
            {
#pragma unroll
                for (int i = 0; i < 16; i++)
                    if (i < ex) w[row + i] = out[i];
            }
            float* mine = buf[g][step & 1][k];
#pragma unroll
            for (int i = 0; i < 16; i++) mine[i] = out[i];
            if (BWD && partials)
            {
                // the final w of the row against the preconditioned residual
                float R[16];
                if (vec)
                {
#pragma unroll
                    for (int v = 0; v < 4; v++)
                    {
                        const float4 q = reinterpret_cast<const float4*>(r + row)[v];
                        R[4*v] = q.x; R[4*v + 1] = q.y; R[4*v + 2] = q.z; R[4*v + 3] = q.w;
                    }
                }
                else
                {
#pragma unroll
                    for (int i = 0; i < 16; i++) R[i] = i < ex ? r[row + i] : 0.f;
                }
                double sr = 0.0;
#pragma unroll
                for (int i = 0; i < 16; i++) sr += (double)out[i]*(double)R[i];
                dotWR += sr;
            }
        }
        __syncthreads();
    }
    if (BWD && partials)
    {
        // one partial per wavefront (= workgroup), fixed butterfly order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) dotWR += __shfl_xor(dotWR, off, 64);
        if (lane == 0)
        {
            partials[blockIdx.x] = dotWR;
            partials[gridDim.x + blockIdx.x] = 0.0;
        }
    }
}

// The whole tile-local apply in one kernel for boxes whose rows are whole
// 16-wide, 16-byte aligned vectors (nx a multiple of 16): the forward walk,
// then the backward walk of the same four tiles by the same wavefront, so
// the backward pass reads the forward result and the tile's coefficients
// while they are still in the caches.  The inputs of the lane's row on the
// next diagonal are loaded while the current diagonal is computed (two
// register sets in turn).  Coefficients of the backward pass: the upper
// arrays, or (symmetric) the lower arrays of the rows (j, k+1), (j+1, k) and
// of the row itself shifted by one cell.  Same per-row arithmetic as
// k_tile_local_f32.
__device__ __forceinline__ float q4(const float4 (&v)[4], int i)
{
    const float4 q = v[i >> 2];
    const int e = i & 3;
    return e == 0 ? q.x : (e == 1 ? q.y : (e == 2 ? q.z : q.w));
}

// 16 coefficients of a row as stored: FP32 (four 16-byte words) or FP16
// (two 16-byte words)
template<class CT> struct CoefRow;
template<> struct CoefRow<float>
{
    float4 v[4];
    __device__ void load(const float* p) { for (int u = 0; u < 4; u++) v[u] = reinterpret_cast<const float4*>(p)[u]; }
    __device__ float at(int i) const { return q4(v, i); }
};
template<> struct CoefRow<__half>
{
    uint4 v[2];
    __device__ void load(const __half* p) { for (int u = 0; u < 2; u++) v[u] = reinterpret_cast<const uint4*>(p)[u]; }
    __device__ float at(int i) const
    {
        const uint4 q = v[i >> 3];
        const int e = (i >> 1) & 3;
        const unsigned int wd = e == 0 ? q.x : (e == 1 ? q.y : (e == 2 ? q.z : q.w));
        return __half2float(__ushort_as_half((unsigned short)(wd >> (16*(i & 1)))));
    }
};

template<class CT>
struct TlRow
{
    float4 A[4], D[4];
    float4 R[4];          // backward pass with the w.r sum: the row of r
    CoefRow<CT> K, J, I;
};

struct TlTile
{
    int ey, ez, nDiag;
    int64_t base, sy, sz;
    bool ok;
};

template<bool BWD>
__device__ __forceinline__ bool tl_row(const TlTile& T, int l16, int step, int& j, int& k)
{
    const int d = BWD ? T.nDiag - 1 - step : step;
    const int kmin = max(0, d - (T.ey - 1));
    k = kmin + l16;
    j = d - k;
    return T.ok && step < T.nDiag && k < T.ez && j >= 0 && j < T.ey;
}

// coefficient arrays of a pass: K, J, I; with SHIFT (symmetric backward)
// read at the rows (j, k+1), (j+1, k) and one cell further along i
// FP16 coefficients are stored as a*2^e (e chosen per matrix so that the
// magnitudes lie in the FP16 normal range); inv = 2^-e turns them back.
// Power-of-two scaling commutes with rounding, so (rD*inv)*h and
// rD*(h*inv) are the same FP32 value: the apply is exactly the one with the
// unscaled rounded coefficients.
template<class CT>
struct TlCoef
{
    const CT *K, *J, *I;
    float inv = 1.f;
};

template<bool BWD, bool SHIFT, bool DOT, class CT>
__device__ __forceinline__ void tl_load(const TlTile& T, int l16, int step, const float* a,
                                        const float* rD, const TlCoef<CT>& C, TlRow<CT>& o,
                                        const float* r)
{
    int j, k;
    const bool act = tl_row<BWD>(T, l16, step, j, k);
    const int64_t row = act ? T.base + T.sy*j + T.sz*k : T.base;
    // neighbour rows inside the tile (else the row itself: value unused)
    const int64_t rowK = (SHIFT && act && k + 1 < T.ez) ? row + T.sz : row;
    const int64_t rowJ = (SHIFT && act && j + 1 < T.ey) ? row + T.sy : row;
#pragma unroll
    for (int v = 0; v < 4; v++)
    {
        o.A[v] = reinterpret_cast<const float4*>(a + row)[v];
        o.D[v] = reinterpret_cast<const float4*>(rD + row)[v];
    }
    o.K.load(C.K + rowK);
    o.J.load(C.J + rowJ);
    o.I.load(C.I + row);
    if constexpr (BWD && DOT)
    {
#pragma unroll
        for (int v = 0; v < 4; v++) o.R[v] = reinterpret_cast<const float4*>(r + row)[v];
    }
}

template<bool BWD, int BZ, bool SHIFT, bool DOT, class CT>
__device__ __forceinline__ void tl_step(const TlTile& T, int l16, int step, const TlRow<CT>& c,
                                        float (*bufg)[BZ + 1][16], float* w, double& dotWR,
                                        double& rsum, float cinv)
{
    constexpr bool scaled = std::is_same<CT, __half>::value;
    int j, k;
    const bool act = tl_row<BWD>(T, l16, step, j, k);
    // every lane computes (an idle lane on its clamped row, results
    // discarded), so the step has no branch around its loads: the inputs
    // loaded ahead are waited for on one path only
    const int kc = act ? k : 0;
    const bool hasK = act && (BWD ? k < T.ez - 1 : k > 0);
    const bool hasJ = act && (BWD ? j < T.ey - 1 : j > 0);
    // the neighbour rows of the previous diagonal, read whole and
    // unconditionally (a missing neighbour's term is selected away:
    // v - 0 == v, the same value as skipping it)
    const int sK = hasK ? (BWD ? k + 1 : k - 1) : kc;
    float4 NK[4], NJ[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
    {
        NK[u] = reinterpret_cast<const float4*>(bufg[(step + 1) & 1][sK])[u];
        NJ[u] = reinterpret_cast<const float4*>(bufg[(step + 1) & 1][kc])[u];
    }
    float out[16];
    float prev = 0.f;
    // forward pass with the sum: t = r - sum coefficient*w(neighbour), the
    // row's w = rD*t; sum w*t (see k_tile_local_vec)
    double st[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 16; q++)
    {
        const int i = BWD ? 15 - q : q;
        const float Di = q4(c.D, i);
        // the coefficient products' factor: rD, times 2^-e for FP16 storage
        const float Dc = scaled ? Di*cinv : Di;
        float v = BWD ? q4(c.A, i) : Di*q4(c.A, i);
        const float tK = Dc*c.K.at(i)*q4(NK, i);
        const float tJ = Dc*c.J.at(i)*q4(NJ, i);
        v -= hasK ? tK : 0.f;
        v -= hasJ ? tJ : 0.f;
        // SHIFT: the upper i coefficient of cell i is the lower one of
        // cell i+1 (cell 15 has no i+1 inside the tile)
        const bool hasI = BWD ? i < 15 : i > 0;
        if (hasI) v -= Dc*c.I.at(SHIFT ? min(i + 1, 15) : i)*prev;
        if constexpr (!BWD && DOT)
        {
            // t scaled by 2^e with the stored coefficients; the row's sum is
            // scaled back below (exact)
            float t = scaled ? q4(c.A, i)/cinv : q4(c.A, i);
            t -= hasK ? c.K.at(i)*q4(NK, i) : 0.f;
            t -= hasJ ? c.J.at(i)*q4(NJ, i) : 0.f;
            if (hasI) t -= c.I.at(i)*prev;
            st[i & 3] += (double)v*(double)t;
        }
        prev = v;
        out[i] = v;
    }
    if constexpr (!BWD && DOT)
    {
        const double t = ((st[0] + st[1]) + (st[2] + st[3]))*(scaled ? (double)cinv : 1.0);
        dotWR += act ? t : 0.0;
        // the row's sum of r (the tile sums of the coarse correction)
        double sr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 16; i++) sr[i & 3] += (double)q4(c.A, i);
        const double u = (sr[0] + sr[1]) + (sr[2] + sr[3]);
        rsum += act ? u : 0.0;
    }
    if constexpr (BWD && DOT)
    {
        // four independent partial sums of the row (fixed order)
        double sr[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int i = 0; i < 16; i++) sr[i & 3] += (double)out[i]*(double)q4(c.R, i);
        const double t = (sr[0] + sr[1]) + (sr[2] + sr[3]);
        dotWR += act ? t : 0.0;
    }
    // the LDS row is written by every lane (an idle lane into the spare
    // row BZ), which keeps the computation above out of the branch
    float* mine = bufg[step & 1][act ? k : BZ];
#pragma unroll
    for (int i = 0; i < 16; i++) mine[i] = out[i];
    if (act)
    {
        const int64_t row = T.base + T.sy*j + T.sz*k;
#pragma unroll
        for (int v = 0; v < 4; v++)
            reinterpret_cast<float4*>(w + row)[v] =
                make_float4(out[4*v], out[4*v + 1], out[4*v + 2], out[4*v + 3]);
    }
    __syncthreads();
}

template<bool BWD, int BZ, bool SHIFT, bool DOT, class CT>
__device__ __forceinline__ void tl_walk(const TlTile& T, int l16, int nD, const float* a,
                                        const float* rD, const TlCoef<CT>& C,
                                        float (*bufg)[BZ + 1][16], float* w, const float* r,
                                        double& dotWR, double& rsum)
{
    TlRow<CT> r0, r1;
    tl_load<BWD, SHIFT, DOT>(T, l16, 0, a, rD, C, r0, r);
    // the scheduling barriers keep each set's loads issued ahead of the
    // step that precedes their use (the compiler would otherwise sink them
    // next to their uses).  Pairs of diagonals in the loop, an odd last one
    // after it: no exit from the middle of the body, whose path would leave
    // the prefetched loads outstanding at the loop head.
    int step = 0;
    for (; step + 1 < nD; step += 2)
    {
        tl_load<BWD, SHIFT, DOT>(T, l16, step + 1, a, rD, C, r1, r);
        __builtin_amdgcn_sched_barrier(0);
        tl_step<BWD, BZ, SHIFT, DOT>(T, l16, step, r0, bufg, w, dotWR, rsum, C.inv);
        tl_load<BWD, SHIFT, DOT>(T, l16, step + 2, a, rD, C, r0, r);
        __builtin_amdgcn_sched_barrier(0);
        tl_step<BWD, BZ, SHIFT, DOT>(T, l16, step + 1, r1, bufg, w, dotWR, rsum, C.inv);
    }
    if (step < nD) tl_step<BWD, BZ, SHIFT, DOT>(T, l16, step, r0, bufg, w, dotWR, rsum, C.inv);
}

// PASSES: 1 forward, 2 backward, 3 forward then backward; TPW tiles per
// wavefront (16 lanes each; the other lanes idle); SYM: the backward pass
// reads the lower coefficient arrays (shifted); DOT: the backward pass sums
// w*r into one partial per workgroup
template<int PASSES, int BZ, int TPW, bool SYM, class CT, bool DOT>
__global__ __launch_bounds__(64)
void k_tile_local_vec(TileView G, const float* x, const float* __restrict__ rD,
                      TlCoef<CT> lo, TlCoef<CT> up, float* w, const DevState* st,
                      const float* __restrict__ r, double* partials, FusedFinalize F,
                      double* tileSum)
{
    if (st && st->done) return;
    __shared__ __align__(16) float buf[TPW][2][BZ + 1][16];
    const int lane = threadIdx.x;
    const int g = min(lane >> 4, TPW - 1), l16 = lane & 15;
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    const int tile = blockIdx.x*TPW + g;
    TlTile T;
    T.ok = (lane >> 4) < TPW && tile < nTiles;
    const int t = T.ok ? tile : 0;
    const int TI = t % G.nTI, TJ = (t/G.nTI) % G.nTJ, TK = t/(G.nTI*G.nTJ);
    T.ey = min(16, G.ny - TJ*16);
    T.ez = min(BZ, G.nz - TK*BZ);
    T.nDiag = T.ey + T.ez - 1;
    T.sy = G.nx;
    T.sz = (int64_t)G.nx*G.ny;
    T.base = TI*16 + T.sy*(TJ*16) + T.sz*(TK*BZ);
    constexpr bool wantDot = (PASSES & 2) && DOT;
    double dotWR = 0.0;
    // the wave walks the longest diagonal count of its tiles
    int nD = T.ok ? T.nDiag : 0;
    nD = max(nD, __shfl_xor(nD, 16, 64));
    nD = max(nD, __shfl_xor(nD, 32, 64));
    nD = __builtin_amdgcn_readfirstlane(nD);
    float (*bufg)[BZ + 1][16] = buf[g];
    // Symmetric, both passes: the sum w*r is taken in the forward pass as
    // sum w'*t, where the forward pass gives w' = rD*t with t = r - L w'
    // (L: the lower coefficients): with M = (D + L) D^-1 (D + L^T) and
    // D = 1/rD, r^T M^-1 r = w'^T D w' = sum w'*t.  The backward pass then
    // does not read r.  Otherwise the backward pass sums w*r.
    constexpr bool dotFwd = wantDot && SYM && PASSES == 3;
    constexpr bool dotBwd = wantDot && !dotFwd;
    double rsum = 0.0;
    if constexpr ((PASSES & 1) != 0)
        tl_walk<false, BZ, false, dotFwd>(T, l16, nD, x, rD, lo, bufg, w, r, dotWR, rsum);
    // (the walk ends on a workgroup barrier: the forward rows are visible)
    if constexpr ((PASSES & 2) != 0)
        tl_walk<true, BZ, SYM, dotBwd>(T, l16, nD, w, rD, up, bufg, w, r, dotWR, rsum);
    if constexpr (dotFwd)
    {
        // tileSum: the sum of r over each tile, the 16 lanes of its group
        // in fixed butterfly order
        if (tileSum)
        {
#pragma unroll
            for (int off = 8; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
            if (l16 == 0 && T.ok) tileSum[tile] = rsum;
        }
    }
    if constexpr (wantDot)
    {
        // one partial per wavefront (= workgroup), fixed butterfly order
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) dotWR += __shfl_xor(dotWR, off, 64);
        partials_arrive(dotWR, 0.0, partials, gridDim.x, blockIdx.x, F);
    }
}

// Quarter-row pipeline of the symmetric tile-local DIC (both passes, one
// 16 x 16 x 16 tile per wavefront, 16 lanes per quarter of an i-row).
//
// The tile's rows are walked by diagonals j + k = d as in k_tile_local_vec,
// but each row is split into four quarters of four cells: lane (q, m),
// lane = q + 4m, processes quarter q of row m of diagonal d = t - q at step
// t.  Its inputs are all from step t - 1: the i - 1 neighbour is the last
// cell of quarter q - 1 of the same row (lane - 1, DPP row shift), the
// (j - 1, k) and (j, k - 1) neighbours are quarter q of rows of diagonal
// d - 1 (LDS, one buffer per diagonal parity).  A step is four cells of
// dependent arithmetic instead of a whole row, and a wavefront owns one tile
// (4096 wavefronts at 256^3: four per SIMD) instead of four, so the walk's
// latency is hidden by the other wavefronts.  The backward pass walks the
// diagonals and the quarters in reverse (lane q processes d = 33 - t - q),
// with the i + 1 neighbour from lane + 1.
//
// Forward (symmetric, with the sum w.r taken as sum w'.t, see
// k_tile_local_vec): t = r - cK w'(k-1) - cJ w'(j-1) - cI w'(i-1),
// w' = rD t; backward: w = w' - rD (cK(k+1) w(k+1) + cJ(j+1) w(j+1) +
// cI(i+1) w(i+1)), the upper coefficients being the lower ones of the
// neighbour cells.  FP16 coefficients are stored scaled by 2^e (scale):
// t is carried scaled by 2^e, rD by 2^-e, and the sum w'.t scaled back.
template<class CT> struct Quarter;
template<> struct Quarter<float>
{
    float4 v;
    __device__ void load(const float* p) { v = *reinterpret_cast<const float4*>(p); }
    __device__ float at(int c) const { return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w)); }
};
template<> struct Quarter<__half>
{
    uint2 v;
    __device__ void load(const __half* p) { v = *reinterpret_cast<const uint2*>(p); }
    __device__ float at(int c) const
    {
        const unsigned int wd = c < 2 ? v.x : v.y;
        return __half2float(__ushort_as_half((unsigned short)(wd >> (16*(c & 1)))));
    }
};
__device__ __forceinline__ float q4at(const float4& v, int c)
{
    return c == 0 ? v.x : (c == 1 ? v.y : (c == 2 ? v.z : v.w));
}
__device__ __forceinline__ float cvt_coef(float c) { return c; }
__device__ __forceinline__ float cvt_coef(__half c) { return __half2float(c); }

struct QRow
{
    int j, k;
    bool act;
    int64_t row;            // first cell of the lane's quarter (a valid cell when !act)
};

// row of diagonal d for the lane's m (k = max(0, d - 15) + m, j = d - k)
__device__ __forceinline__ QRow q_row(int d, int m, int ey, int ez, int64_t base, int64_t sy,
                                      int64_t sz)
{
    QRow R;
    R.k = max(0, d - 15) + m;
    R.j = d - R.k;
    R.act = d >= 0 && d <= 30 && R.k < ez && R.j >= 0 && R.j < ey;
    R.row = R.act ? base + sy*R.j + sz*R.k : base;
    return R;
}

template<class CT>
struct QIn
{
    float4 X, D;
    Quarter<CT> K, J, I;
    float I4;               // backward: the lower i coefficient of the next quarter's first cell
};

template<bool BWD, class CT>
__device__ __forceinline__ void q_load(const QRow& R, int q, int ey, int ez, int64_t sy,
                                       int64_t sz, const float* x, const float* rD,
                                       const TlCoef<CT>& C, QIn<CT>& o)
{
    o.X = *reinterpret_cast<const float4*>(x + R.row);
    o.D = *reinterpret_cast<const float4*>(rD + R.row);
    if constexpr (!BWD)
    {
        o.K.load(C.K + R.row);
        o.J.load(C.J + R.row);
        o.I.load(C.I + R.row);
    }
    else
    {
        // upper coefficients: the lower ones of (j, k+1), (j+1, k) and of
        // the cells one further along i (read where the neighbour exists)
        o.K.load(C.K + (R.act && R.k + 1 < ez ? R.row + sz : R.row));
        o.J.load(C.J + (R.act && R.j + 1 < ey ? R.row + sy : R.row));
        o.I.load(C.I + R.row);
        o.I4 = q < 3 ? cvt_coef(C.I[R.row + 4]) : 0.f;
    }
}

template<class CT, bool DOT>
__global__ __launch_bounds__(64)
void k_tile_dic_q(TileView G, const float* x, const float* __restrict__ rD, TlCoef<CT> C,
                  float scale, float* w, const DevState* st, double* partials, FusedFinalize F,
                  double* tileSum)
{
    if (st && st->done) return;
    constexpr bool scaled = std::is_same<CT, __half>::value;
    __shared__ __align__(16) float buf[2][17][16];
    const int lane = threadIdx.x;
    const int q = lane & 3, m = lane >> 2;
    // XCD-aware order (see k_tile_dic_lds): tiles adjacent along i, whose
    // rows share cache lines, run together on one XCD's L2
    const int nT = G.nTI*G.nTJ*G.nTK;
    const int tile = (nT & 7) == 0 ? (int)(blockIdx.x & 7)*(nT >> 3) + (int)(blockIdx.x >> 3)
                                   : (int)blockIdx.x;
    const int TI = tile % G.nTI, TJ = (tile/G.nTI) % G.nTJ, TK = tile/(G.nTI*G.nTJ);
    const int ey = min(16, G.ny - TJ*16), ez = min(16, G.nz - TK*16);
    const int64_t sy = G.nx, sz = (int64_t)G.nx*G.ny;
    const int64_t base = TI*16 + sy*(TJ*16) + sz*(TK*16) + 4*q;
    const float inv = C.inv;
    double dotWR = 0.0, rsum = 0.0;

    // ---- forward ----
    {
        QIn<CT> in0, in1;
        QRow R0 = q_row(-q, m, ey, ez, base, sy, sz), R1;
        q_load<false>(R0, q, ey, ez, sy, sz, x, rD, C, in0);
        float last = 0.f;       // this lane's last cell of the previous step
        auto step = [&](int t, const QRow& R, const QIn<CT>& in) {
            const int d = t - q;
            const bool hasK = R.act && R.k > 0;
            const bool hasJ = R.act && R.j > 0;
            const float4 NK = *reinterpret_cast<const float4*>(&buf[(d + 1) & 1][hasK ? R.k - 1 : 16][4*q]);
            const float4 NJ = *reinterpret_cast<const float4*>(&buf[(d + 1) & 1][hasJ ? R.k : 16][4*q]);
            float prev = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(last),
                                        __float_as_int(last), 0x111, 0xf, 0xf, false));
            prev = q > 0 ? prev : 0.f;
            float out[4];
            float dp = 0.f, rs = 0.f;
#pragma unroll
            for (int c = 0; c < 4; c++)
            {
                const float xc = q4at(in.X, c);
                float ts = scaled ? xc*scale : xc;
                ts = __builtin_fmaf(-in.K.at(c), hasK ? q4at(NK, c) : 0.f, ts);
                ts = __builtin_fmaf(-in.J.at(c), hasJ ? q4at(NJ, c) : 0.f, ts);
                ts = __builtin_fmaf(-in.I.at(c), prev, ts);
                const float Dc = scaled ? q4at(in.D, c)*inv : q4at(in.D, c);
                const float v = Dc*ts;
                if constexpr (DOT) dp = __builtin_fmaf(v, ts, dp);
                rs += xc;
                prev = v;
                out[c] = v;
            }
            last = out[3];
            const float4 o4 = make_float4(out[0], out[1], out[2], out[3]);
            *reinterpret_cast<float4*>(&buf[d & 1][R.act ? R.k : 16][4*q]) = o4;
            if (R.act)
            {
                *reinterpret_cast<float4*>(w + R.row) = o4;
                if constexpr (DOT)
                {
                    dotWR += (double)dp;
                    rsum += (double)rs;
                }
            }
            __syncthreads();
        };
        // 34 steps: two per loop pass, the next step's inputs loaded ahead
        for (int t = 0; t < 34; t += 2)
        {
            R1 = q_row(t + 1 - q, m, ey, ez, base, sy, sz);
            q_load<false>(R1, q, ey, ez, sy, sz, x, rD, C, in1);
            __builtin_amdgcn_sched_barrier(0);
            step(t, R0, in0);
            R0 = q_row(t + 2 - q, m, ey, ez, base, sy, sz);
            if (t + 2 < 34) q_load<false>(R0, q, ey, ez, sy, sz, x, rD, C, in0);
            __builtin_amdgcn_sched_barrier(0);
            step(t + 1, R1, in1);
        }
    }
    // ---- backward (w in place: each lane re-reads what it wrote) ----
    {
        QIn<CT> in0, in1;
        QRow R0 = q_row(33 - q, m, ey, ez, base, sy, sz), R1;
        q_load<true>(R0, q, ey, ez, sy, sz, w, rD, C, in0);
        float last = 0.f;
        auto step = [&](int t, const QRow& R, const QIn<CT>& in) {
            const int d = 33 - t - q;
            const bool hasK = R.act && R.k + 1 < ez;
            const bool hasJ = R.act && R.j + 1 < ey;
            const float4 NK = *reinterpret_cast<const float4*>(&buf[(d + 1) & 1][hasK ? R.k + 1 : 16][4*q]);
            const float4 NJ = *reinterpret_cast<const float4*>(&buf[(d + 1) & 1][hasJ ? R.k : 16][4*q]);
            float next = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(last),
                                        __float_as_int(last), 0x101, 0xf, 0xf, false));
            float out[4];
#pragma unroll
            for (int cc = 0; cc < 4; cc++)
            {
                const int c = 3 - cc;
                const float cI = c < 3 ? in.I.at(c + 1) : in.I4;
                // cell 15 of the row has no i + 1 neighbour (cI = 0 there
                // only when q = 3, where I4 is 0)
                float acc = in.K.at(c)*(hasK ? q4at(NK, c) : 0.f);
                acc = __builtin_fmaf(in.J.at(c), hasJ ? q4at(NJ, c) : 0.f, acc);
                acc = __builtin_fmaf(cI, (c < 3 || q < 3) ? next : 0.f, acc);
                const float Dc = scaled ? q4at(in.D, c)*inv : q4at(in.D, c);
                const float v = __builtin_fmaf(-Dc, acc, q4at(in.X, c));
                next = v;
                out[c] = v;
            }
            last = out[0];
            const float4 o4 = make_float4(out[0], out[1], out[2], out[3]);
            *reinterpret_cast<float4*>(&buf[d & 1][R.act ? R.k : 16][4*q]) = o4;
            if (R.act) *reinterpret_cast<float4*>(w + R.row) = o4;
            __syncthreads();
        };
        for (int t = 0; t < 34; t += 2)
        {
            R1 = q_row(33 - (t + 1) - q, m, ey, ez, base, sy, sz);
            q_load<true>(R1, q, ey, ez, sy, sz, w, rD, C, in1);
            __builtin_amdgcn_sched_barrier(0);
            step(t, R0, in0);
            R0 = q_row(33 - (t + 2) - q, m, ey, ez, base, sy, sz);
            if (t + 2 < 34) q_load<true>(R0, q, ey, ez, sy, sz, w, rD, C, in0);
            __builtin_amdgcn_sched_barrier(0);
            step(t + 1, R1, in1);
        }
    }
    if constexpr (DOT)
    {
        if (tileSum)
        {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
            if (lane == 0) tileSum[tile] = rsum;
        }
        dotWR *= scaled ? (double)inv : 1.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) dotWR += __shfl_xor(dotWR, off, 64);
        partials_arrive(dotWR, 0.0, partials, gridDim.x, blockIdx.x, F);
    }
}

// The symmetric tile-local DIC with the whole 16 x 16 x 16 tile staged in
// LDS (r, rD and the FP16 coefficients, 59 KB): one wavefront per tile
// stages it with whole-row vector loads, runs both triangular passes out of
// LDS and writes w once, so every input is read from HBM once and w written
// once.
//
// Lane (j, q) = (lane & 15, lane >> 4) owns quarter q (cells 4q .. 4q+3) of
// the 16 rows (j, k), k = 0 .. 15.  Forward step t takes k = t - j - q: the
// (j, k-1) neighbour is the lane's own result of step t-1, the (j-1, k) one
// that of lane j-1 at step t-1 (DPP row shift, same 16-lane row) and the i-1
// neighbour, the last cell of quarter q-1, that of lane j + 16(q-1) (a lane
// shuffle).  The backward pass mirrors it (k = 33 - t - j - q).  A lane's own
// r, rD, coefficients and forward results are read from and written to LDS
// at its own addresses; only the read-only coefficients of the upper
// neighbours are another lane's, so the passes need no barrier.  Planes are
// 268 cells apart in LDS: the 16 lanes of a step then fall on distinct banks.
constexpr int TL_PLANE = 268;
constexpr int TL_LDS_CELLS = 15*TL_PLANE + 256 + 20;     // a multiple of 8
// FP16 coefficients: pairs of planes (one direct-to-LDS load each) 520 cells
// apart, the layout of the fewest bank conflicts among those the loads allow
constexpr int TL_HPAIR = 520;
constexpr int TL_HCELLS = 8*TL_HPAIR;
static_assert(TL_LDS_CELLS % 8 == 0 && TL_PLANE % 4 == 0 && TL_HPAIR % 8 == 0,
              "LDS tile layout alignment");

__device__ inline int tl_lds(int l)   // tile cell l = i + 16 (j + 16 k) -> LDS cell (FP32)
{
    return (l & 255) + TL_PLANE*(l >> 8);
}

__device__ inline int tl_hlds(int j, int k, int q)   // quarter q of row (j, k) -> FP16 LDS cell
{
    return 4*q + 16*j + 256*(k & 1) + TL_HPAIR*(k >> 1);
}

// 16 bytes per lane from global memory straight into LDS: lane L writes
// bytes [16 L, 16 L + 16) after the wave-uniform LDS address `l`
__device__ inline void tl_dma16(const void* g, void* l)
{
    __builtin_amdgcn_global_load_lds(
        (const __attribute__((address_space(1))) void*)g,
        (__attribute__((address_space(3))) void*)l, 16, 0, 0);
}

__device__ inline float4 tl_half4(const __half* a, int c)   // cells c .. c+3, c % 4 == 0
{
    const uint2 v = *reinterpret_cast<const uint2*>(a + c);
    const __half2 lo = *reinterpret_cast<const __half2*>(&v.x);
    const __half2 hi = *reinterpret_cast<const __half2*>(&v.y);
    return make_float4(__low2float(lo), __high2float(lo), __low2float(hi), __high2float(hi));
}

__device__ inline float tl_dpp_from_lower(float v)    // lane j-1 of the 16-lane row; 0 at j = 0
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf,
                                                      false));
}

__device__ inline float tl_dpp_from_upper(float v)    // lane j+1 of the 16-lane row; 0 at j = 15
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf,
                                                      false));
}

// the value of the previous (dn) and of the next (up) 16-lane row's lane in
// the same position, 0 in row 0 (dn) and row 3 (up): VALU lane swaps, no
// LDS round trip (permlane16_swap exchanges rows 0<->1 and 2<->3 between
// its operands, permlane32_swap the two 32-lane halves)
__device__ inline float2 tl_rows_dn_up(float v)
{
    const auto a = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v),
                                                    false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(a[1], a[1], false, false);
    const auto c = __builtin_amdgcn_permlane32_swap(a[0], a[0], false, false);
    const int row = threadIdx.x >> 4;
    const float dn = row == 0 ? 0.f : __uint_as_float(row == 2 ? b[0] : a[0]);
    const float up = row == 3 ? 0.f : __uint_as_float(row == 1 ? c[1] : a[1]);
    return make_float2(dn, up);
}

template<bool DOT>
__global__ __launch_bounds__(64)
void k_tile_dic_lds(TileView G, const float* x, const float* __restrict__ rD,
                    TlCoef<__half> C, float scale, float* w, const DevState* st,
                    double* partials, FusedFinalize F, double* tileSum)
{
    if (st && st->done) return;
    extern __shared__ __align__(16) unsigned char lds[];
    float* X = reinterpret_cast<float*>(lds);
    float* Dv = X + TL_LDS_CELLS;
    __half* cK = reinterpret_cast<__half*>(Dv + TL_LDS_CELLS);
    __half* cJ = cK + TL_HCELLS;
    __half* cI = cJ + TL_HCELLS;
    const int lane = threadIdx.x;
    // XCD-aware order: workgroups go round robin over the 8 XCDs, so XCD x
    // is given the x-th eighth of the tiles in order -- tiles adjacent along
    // i, whose rows share cache lines, then run together on one L2
    const int nT = G.nTI*G.nTJ*G.nTK;
    const int tile = (nT & 7) == 0 ? (int)(blockIdx.x & 7)*(nT >> 3) + (int)(blockIdx.x >> 3)
                                   : (int)blockIdx.x;
    const int TI = tile % G.nTI, TJ = (tile/G.nTI) % G.nTJ, TK = tile/(G.nTI*G.nTJ);
    const int ey = min(16, G.ny - TJ*16), ez = min(16, G.nz - TK*16);
    const int64_t sy = G.nx, sz = (int64_t)G.nx*G.ny;
    const int64_t base = TI*16 + sy*(TJ*16) + sz*(TK*16);
    // ---- stage with direct-to-LDS loads: one instruction per plane of r
    // and rD (lane L: row L/4, quarter L%4), one per pair of planes of each
    // FP16 coefficient array (lane L: plane L/32, row (L%32)/2, half-row
    // L%2).  Rows and planes outside the mesh load the nearest ones inside
    // it, so every slot holds finite values: the slots outside the mesh are
    // only ever multiplied by a zero neighbour value.
    {
        const int jr = min(lane >> 2, ey - 1);
        const int64_t gr = base + 4*(lane & 3) + sy*jr;
#pragma unroll
        for (int k = 0; k < 16; k++)
        {
            const int64_t g = gr + sz*min(k, ez - 1);
            tl_dma16(x + g, X + TL_PLANE*k);
            tl_dma16(rD + g, Dv + TL_PLANE*k);
        }
        const int jh = min((lane & 31) >> 1, ey - 1);
        const int64_t gh = base + 8*(lane & 1) + sy*jh;
#pragma unroll
        for (int p = 0; p < 8; p++)
        {
            const int64_t g = gh + sz*min(2*p + (lane >> 5), ez - 1);
            tl_dma16(C.K + g, cK + TL_HPAIR*p);
            tl_dma16(C.J + g, cJ + TL_HPAIR*p);
            tl_dma16(C.I + g, cI + TL_HPAIR*p);
        }
        // vmcnt(0): the direct-to-LDS loads have landed
        __builtin_amdgcn_s_waitcnt(0x0F70);
    }
    __syncthreads();
    const int j = lane & 15, q = lane >> 4;
    const bool rowIn = j < ey;
    const float inv = C.inv;
    const float4 zero4 = make_float4(0.f, 0.f, 0.f, 0.f);
    auto h4 = [](const __half* a, int h) { return tl_half4(a, h); };
    double dotWR = 0.0, rsum = 0.0;
    // ---- forward ----
    // po, the lane's result of its previous step, is zero until its first
    // row (k = 0): the (j, k-1) value needs no mask, the (j-1, k) one comes
    // from lane j-1 (0 at j = 0) and the i-1 one from row q-1 (0 at q = 0)
    {
        float4 po = zero4;
        for (int t = 0; t < 34; t++)
        {
            const int k = t - j - q;
            const bool act = rowIn && k >= 0 && k < ez;
            float4 nj;
            nj.x = tl_dpp_from_lower(po.x);
            nj.y = tl_dpp_from_lower(po.y);
            nj.z = tl_dpp_from_lower(po.z);
            nj.w = tl_dpp_from_lower(po.w);
            float prev = tl_rows_dn_up(po.w).x;
            if (act)
            {
                const int c0 = 4*q + 16*j + TL_PLANE*k, h0 = tl_hlds(j, k, q);
                const float4 xr = *reinterpret_cast<const float4*>(X + c0);
                const float4 dr = *reinterpret_cast<const float4*>(Dv + c0);
                const float4 ck = h4(cK, h0), cj = h4(cJ, h0), ci = h4(cI, h0);
                const float4 nk = po;
                float dp = 0.f;
                auto fwd = [&](float xv, float dv, float kv, float jv, float a, float b, float c) {
                    float ts = xv*scale;
                    ts = __builtin_fmaf(-a, kv, ts);
                    ts = __builtin_fmaf(-b, jv, ts);
                    ts = __builtin_fmaf(-c, prev, ts);
                    const float v = (dv*inv)*ts;
                    if constexpr (DOT) dp = __builtin_fmaf(v, ts, dp);
                    prev = v;
                    return v;
                };
                float4 o;
                o.x = fwd(xr.x, dr.x, nk.x, nj.x, ck.x, cj.x, ci.x);
                o.y = fwd(xr.y, dr.y, nk.y, nj.y, ck.y, cj.y, ci.y);
                o.z = fwd(xr.z, dr.z, nk.z, nj.z, ck.z, cj.z, ci.z);
                o.w = fwd(xr.w, dr.w, nk.w, nj.w, ck.w, cj.w, ci.w);
                *reinterpret_cast<float4*>(X + c0) = o;
                po = o;
                if constexpr (DOT)
                {
                    dotWR += (double)dp;
                    rsum += (double)((xr.x + xr.y) + (xr.z + xr.w));
                }
            }
        }
    }
    // ---- backward (in place) ----
    // po is zero until the lane's first row (k = ez-1); lane j+1 of a row
    // outside the mesh (or j = 15) gives 0, as does row q+1 at q = 3
    {
        float4 po = zero4;
        for (int t = 0; t < 34; t++)
        {
            const int k = 33 - t - j - q;
            const bool act = rowIn && k >= 0 && k < ez;
            float4 nj;
            nj.x = tl_dpp_from_upper(po.x);
            nj.y = tl_dpp_from_upper(po.y);
            nj.z = tl_dpp_from_upper(po.z);
            nj.w = tl_dpp_from_upper(po.w);
            float next = tl_rows_dn_up(po.x).y;
            if (act)
            {
                const int c0 = 4*q + 16*j + TL_PLANE*k, h0 = tl_hlds(j, k, q);
                const float4 xr = *reinterpret_cast<const float4*>(X + c0);
                const float4 dr = *reinterpret_cast<const float4*>(Dv + c0);
                // upper coefficients: the lower ones of (j, k+1), (j+1, k) and
                // of the next cell along i (slots past the tile's end hold
                // finite values whose neighbour value is zero)
                const float4 ck = h4(cK, tl_hlds(j, min(k + 1, 15), q));
                const float4 cj = h4(cJ, h0 + (j < 15 ? 16 : 0));
                const float4 ci = h4(cI, h0);
                const float ci4 = __half2float(cI[h0 + (q < 3 ? 4 : 0)]);
                const float4 nk = po;
                auto bwd = [&](float xv, float dv, float kv, float jv, float a, float b, float c) {
                    float acc = a*kv;
                    acc = __builtin_fmaf(b, jv, acc);
                    acc = __builtin_fmaf(c, next, acc);
                    const float v = __builtin_fmaf(-(dv*inv), acc, xv);
                    next = v;
                    return v;
                };
                float4 o;
                o.w = bwd(xr.w, dr.w, nk.w, nj.w, ck.w, cj.w, ci4);
                o.z = bwd(xr.z, dr.z, nk.z, nj.z, ck.z, cj.z, ci.w);
                o.y = bwd(xr.y, dr.y, nk.y, nj.y, ck.y, cj.y, ci.z);
                o.x = bwd(xr.x, dr.x, nk.x, nj.x, ck.x, cj.x, ci.y);
                *reinterpret_cast<float4*>(X + c0) = o;
                po = o;
            }
        }
    }
    __syncthreads();
    // ---- write w (whole rows) ----
#pragma unroll 8
    for (int u = lane; u < 1024; u += 64)
    {
        const int l = 4*u, jj = (l >> 4) & 15, kk = l >> 8;
        if (jj < ey && kk < ez)
            *reinterpret_cast<float4*>(w + base + (l & 15) + sy*jj + sz*kk) =
                *reinterpret_cast<const float4*>(X + tl_lds(l));
    }
    if constexpr (DOT)
    {
        if (tileSum)
        {
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) rsum += __shfl_xor(rsum, off, 64);
            if (lane == 0) tileSum[tile] = rsum;
        }
        dotWR *= (double)inv;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) dotWR += __shfl_xor(dotWR, off, 64);
        partials_arrive(dotWR, 0.0, partials, gridDim.x, blockIdx.x, F);
    }
}

// launch with the w*r sum compiled in when partials are given
template<int PASSES, int BZ, bool SYM, class CT>
static void tlv_launch(dim3 grid, hipStream_t s, const TileView& G, const float* x,
                       const float* rD, const TlCoef<CT>& lo, const TlCoef<CT>& up, float* w,
                       const DevState* st, const float* r, double* partials,
                       const FusedFinalize& F, double* tileSum = nullptr)
{
    if (partials)
        hipLaunchKernelGGL((k_tile_local_vec<PASSES, BZ, 4, SYM, CT, true>), grid, dim3(64), 0,
                           s, G, x, rD, lo, up, w, st, r, partials, F, tileSum);
    else
        hipLaunchKernelGGL((k_tile_local_vec<PASSES, BZ, 4, SYM, CT, false>), grid, dim3(64), 0,
                           s, G, x, rD, lo, up, w, st, r, partials, F, tileSum);
}

int32_t launch_tile_local_apply_f32(const TileView& G, const float* x, const float* rD,
                                    const float* const dir[6], bool symmetric, float* w,
                                    const DevState* st, hipStream_t s, double* partials,
                                    const __half* const half3[3], const FusedFinalize& F,
                                    double* tileSum, float halfScale)
{
    if (G.bx != 16 || G.by != 16 || G.bz != TileBZ<float>::pre || (G.nx & 15))
        return -1;
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    const dim3 grid((nTiles + 3)/4);
    constexpr int BZ = TileBZ<float>::pre;
    // symmetric: the row walk (default; it reads the tile's inputs again in
    // the backward pass), the tile staged in LDS (FOAMHIP_TILE_APPLY=lds:
    // inputs read once, but two tiles per CU leave its latency-bound steps
    // on one wavefront per SIMD) or the quarter-row pipeline from HBM
    // (FOAMHIP_TILE_APPLY=quarter; its per-lane quarter rows touch four
    // times the cache lines)
    // (read at every launch, so a process can switch between them).  By
    // default the staged tile up to 512 tiles (all tiles in one round at two
    // per CU: its shorter per-tile latency wins -- a serving sample of 16^3
    // to 128^3 boxes: 196 against 180 solves/s, p50 15 against 31 ms), the
    // row walk above (256^3: 149 against 167 us)
    const char* env = std::getenv("FOAMHIP_TILE_APPLY");
    const int mode = env && std::strcmp(env, "lds") == 0 ? 0
                   : env && std::strcmp(env, "quarter") == 0 ? 1
                   : env && std::strcmp(env, "rows") == 0 ? 2
                   : nTiles <= 512 ? 0 : 2;
    // the staged tile (FP16 coefficients): 59 KB of LDS per workgroup, taken
    // when the device allows that much
    const size_t ldsBytes = (size_t)TL_LDS_CELLS*(4 + 4) + (size_t)TL_HCELLS*3*sizeof(__half);
    static const size_t ldsMax = [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
            return (size_t)0;
        return (size_t)v;
    }();
    if (symmetric && half3 && BZ == 16 && mode == 0 && ldsBytes <= ldsMax)
    {
        const dim3 gq((unsigned)nTiles);
        const TlCoef<__half> h{half3[0], half3[1], half3[2], 1.f/halfScale};
        if (partials)
            hipLaunchKernelGGL((k_tile_dic_lds<true>), gq, dim3(64), ldsBytes, s, G, x, rD, h,
                               halfScale, w, st, partials, F, tileSum);
        else
            hipLaunchKernelGGL((k_tile_dic_lds<false>), gq, dim3(64), ldsBytes, s, G, x, rD, h,
                               halfScale, w, st, partials, F, tileSum);
        return (int32_t)gq.x;
    }
    if (symmetric && BZ == 16 && mode == 1)
    {
        // quarter-row pipeline: one tile per wavefront
        const dim3 gq((unsigned)nTiles);
        if (half3)
        {
            const TlCoef<__half> h{half3[0], half3[1], half3[2], 1.f/halfScale};
            if (partials)
                hipLaunchKernelGGL((k_tile_dic_q<__half, true>), gq, dim3(64), 0, s, G, x, rD, h,
                                   halfScale, w, st, partials, F, tileSum);
            else
                hipLaunchKernelGGL((k_tile_dic_q<__half, false>), gq, dim3(64), 0, s, G, x, rD,
                                   h, halfScale, w, st, partials, F, tileSum);
        }
        else
        {
            const TlCoef<float> c{dir[0], dir[1], dir[2]};
            if (partials)
                hipLaunchKernelGGL((k_tile_dic_q<float, true>), gq, dim3(64), 0, s, G, x, rD, c,
                                   1.f, w, st, partials, F, tileSum);
            else
                hipLaunchKernelGGL((k_tile_dic_q<float, false>), gq, dim3(64), 0, s, G, x, rD, c,
                                   1.f, w, st, partials, F, tileSum);
        }
        return (int32_t)gq.x;
    }
    if (symmetric && half3)
    {
        const TlCoef<__half> h{half3[0], half3[1], half3[2], 1.f/halfScale};
        tlv_launch<3, BZ, true>(grid, s, G, x, rD, h, h, w, st, x, partials, F, tileSum);
        return (int32_t)grid.x;
    }
    const TlCoef<float> lo{dir[0], dir[1], dir[2]};
    const TlCoef<float> up = symmetric ? TlCoef<float>{dir[0], dir[1], dir[2]}
                                       : TlCoef<float>{dir[5], dir[4], dir[3]};
    if (symmetric)
        tlv_launch<3, BZ, true>(grid, s, G, x, rD, lo, up, w, st, x, partials, F, tileSum);
    else
        tlv_launch<3, BZ, false>(grid, s, G, x, rD, lo, up, w, st, x, partials, F);
    return (int32_t)grid.x;
}

// FP16 copies of the three lower coefficient arrays of a box, scaled by
// 2^e (scale)
__global__ void k_pack_half(int64_t n, const float* __restrict__ a, const float* __restrict__ b,
                            const float* __restrict__ c, __half* __restrict__ ha,
                            __half* __restrict__ hb, __half* __restrict__ hc, float scale)
{
    const int64_t i = (int64_t)blockIdx.x*TL_B + threadIdx.x;
    if (i >= n) return;
    ha[i] = __float2half_rn(a[i]*scale);
    hb[i] = __float2half_rn(b[i]*scale);
    hc[i] = __float2half_rn(c[i]*scale);
}

// values rounded to FP16 precision at the storage scale 2^e, kept in FP32
// and unscaled (exact: power-of-two factors)
__global__ void k_round_half(int64_t n, const float* __restrict__ a, float* __restrict__ b,
                             float scale)
{
    const int64_t i = (int64_t)blockIdx.x*TL_B + threadIdx.x;
    if (i < n) b[i] = __half2float(__float2half_rn(a[i]*scale))/scale;
}

// largest and smallest non-zero |a| as FP32 bit patterns (non-negative
// floats order as unsigned integers): range[0] max (start 0), range[1] min
// (start 0xffffffff)
__global__ void k_abs_range(int64_t n, const float* __restrict__ a, unsigned int* range)
{
    unsigned int mx = 0u, mn = 0xffffffffu;
    for (int64_t i = (int64_t)blockIdx.x*TL_B + threadIdx.x; i < n; i += (int64_t)gridDim.x*TL_B)
    {
        const unsigned int u = __float_as_uint(fabsf(a[i]));
        if (u != 0u)
        {
            mx = max(mx, u);
            mn = min(mn, u);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
    {
        mx = max(mx, (unsigned int)__shfl_xor((int)mx, off, 64));
        mn = min(mn, (unsigned int)__shfl_xor((int)mn, off, 64));
    }
    if ((threadIdx.x & 63) == 0)
    {
        atomicMax(&range[0], mx);
        atomicMin(&range[1], mn);
    }
}

void launch_round_half(int64_t n, const float* a, float* b, float scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_round_half, dim3((unsigned)((n + TL_B - 1)/TL_B)), dim3(TL_B), 0, s, n,
                       a, b, scale);
}

void launch_pack_half(int64_t n, const float* const src[3], __half* const dst[3], float scale,
                      hipStream_t s)
{
    hipLaunchKernelGGL(k_pack_half, dim3((unsigned)((n + TL_B - 1)/TL_B)), dim3(TL_B), 0, s, n,
                       src[0], src[1], src[2], dst[0], dst[1], dst[2], scale);
}

void launch_abs_range(int64_t n, const float* a, unsigned int* range, hipStream_t s)
{
    const int64_t blocks = std::min<int64_t>((n + TL_B - 1)/TL_B, 2048);
    hipLaunchKernelGGL(k_abs_range, dim3((unsigned)std::max<int64_t>(blocks, 1)), dim3(TL_B), 0,
                       s, n, a, range);
}

// Box fields with rows padded to a multiple of 16 cells (the layout of the
// fused tile-local kernels): dst[i + nxp (j + ny k)] = src[i + nx (j + ny k)]
// for i < nx, 0 in the padding (pad), and back (unpad)
__global__ void k_pad_rows(int32_t nx, int32_t nxp, int64_t np, const float* __restrict__ src,
                           float* __restrict__ dst)
{
    const int64_t q = (int64_t)blockIdx.x*TL_B + threadIdx.x;
    if (q >= np) return;
    const int32_t i = (int32_t)(q % nxp);
    const int64_t row = q/nxp;
    dst[q] = i < nx ? src[i + row*nx] : 0.f;
}

__global__ void k_unpad_rows(int32_t nx, int32_t nxp, int64_t n, const float* __restrict__ src,
                             float* __restrict__ dst)
{
    const int64_t c = (int64_t)blockIdx.x*TL_B + threadIdx.x;
    if (c >= n) return;
    const int32_t i = (int32_t)(c % nx);
    dst[c] = src[i + (c/nx)*nxp];
}

void launch_pad_rows(int32_t nx, int32_t nxp, int64_t rows, const float* src, float* dst,
                     hipStream_t s)
{
    const int64_t np = (int64_t)nxp*rows;
    if (np > 0)
        hipLaunchKernelGGL(k_pad_rows, dim3((unsigned)((np + TL_B - 1)/TL_B)), dim3(TL_B), 0, s,
                           nx, nxp, np, src, dst);
}

void launch_unpad_rows(int32_t nx, int32_t nxp, int64_t rows, const float* src, float* dst,
                       hipStream_t s)
{
    const int64_t n = (int64_t)nx*rows;
    if (n > 0)
        hipLaunchKernelGGL(k_unpad_rows, dim3((unsigned)((n + TL_B - 1)/TL_B)), dim3(TL_B), 0, s,
                           nx, nxp, n, src, dst);
}

int32_t launch_tile_local_f32(bool backward, const TileView& G, const float* a, const float* rD,
                              const float* cK, const float* cJ, const float* cI, float* w,
                              const DevState* st, hipStream_t s, const float* r,
                              double* partials)
{
    if (G.bx != 16 || G.by != 16 || G.bz != TileBZ<float>::pre)
        fail(FH_ERR_INVALID, "tile shape does not match the compiled local kernels");
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    const dim3 grid((nTiles + 3)/4);
    if ((G.nx & 15) == 0)
    {
        const TlCoef<float> c{cK, cJ, cI};
        if (backward)
            tlv_launch<2, TileBZ<float>::pre, false>(grid, s, G, a, rD, c, c, w, st, r, partials,
                                                     FusedFinalize());
        else
            hipLaunchKernelGGL((k_tile_local_vec<1, TileBZ<float>::pre, 4, false, float, false>),
                               grid, dim3(64), 0, s, G, a, rD, c, c, w, st, nullptr, nullptr,
                               FusedFinalize(), nullptr);
        return (int32_t)grid.x;
    }
    if (backward)
        hipLaunchKernelGGL((k_tile_local_f32<true, TileBZ<float>::pre>), grid, dim3(64), 0, s,
                           G, a, rD, cK, cJ, cI, w, st, r, partials);
    else
        hipLaunchKernelGGL((k_tile_local_f32<false, TileBZ<float>::pre>), grid, dim3(64), 0, s,
                           G, a, rD, cK, cJ, cI, w, st, nullptr, nullptr);
    return (int32_t)grid.x;
}

} // namespace fh
