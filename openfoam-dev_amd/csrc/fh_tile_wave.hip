// One-wavefront tiled DIC / DILU sweeps for hex boxes (the global
// preconditioner passes of the FP64 and FP32 Krylov solvers).
//
// The recurrences are those of fh_tile_sweep.hip (DICPreconditioner.C:
// 87-123, DILUPreconditioner.C:88-135): cell (i,j,k) waits for its lower
// neighbours (forward) or upper neighbours (backward), and each row keeps
// the reference face order, so the results are bitwise those of the
// sequential loops.  What differs is the shape of the work: a tile is
// 8 x 8 x 16 cells and one wavefront owns it -- lane (i,j) owns the pencil
// along k and the wave walks the tile's 30 hyperplanes i+j+k = s with no
// workgroup barrier, exchanging the previous hyperplane's values with lane
// shuffles.  The walk of a tile is then a few microseconds, close to the
// latency of the cross-tile hand-off (completion flag + halo loads), and a
// CU holds several tiles so that the staging of the next tiles overlaps the
// walk of the current ones.
//
// Tiles are taken by ticket in wavefront order (reversed for backward), a
// tile waits for the completion flags of its (up to three) upstream tiles,
// reads their boundary values with device-coherent loads and publishes its
// own completion after its stores have completed -- as in fh_tile_sweep.hip.
#include "fh_tile_sweep.h"
#include "fh_finalize_dev.h"

namespace fh
{

#define WAVE_BX 8
#define WAVE_BZ 16
#define WAVE_SPIN_LIMIT (1u << 20)

template<class T>
__device__ __forceinline__ T wv_ld_coh(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<class T>
__device__ __forceinline__ void wv_st_coh(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// value of lane src (any lane may ask for any lane)
__device__ __forceinline__ double wv_shfl(double v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ float wv_shfl(float v, int src) { return __shfl(v, src, 64); }

// value of the neighbouring lane within a 16-lane row (DPP row shift by
// one: lane - 1 for UP, lane + 1 otherwise; the row's end lane reads its
// own value -- only lanes whose neighbour is in the row use the result)
template<bool UP>
__device__ __forceinline__ int wv_row_shift(int v)
{
    return __builtin_amdgcn_update_dpp(v, v, UP ? 0x111 : 0x101, 0xf, 0xf, false);
}
template<bool UP>
__device__ __forceinline__ float wv_row_shift(float v)
{
    return __int_as_float(wv_row_shift<UP>(__float_as_int(v)));
}
template<bool UP>
__device__ __forceinline__ double wv_row_shift(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = wv_row_shift<UP>((int)(b & 0xffffffffll));
    const int hi = wv_row_shift<UP>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// dat.dir[0..2]: the pass's coefficients (k, j, i) packed per cell; with
// dat.dirShift read at the neighbour (symmetric backward).  Backward with
// dat.dotR: also the sum w*dotR of the tile into dat.partials[tile].
template<class T, bool BWD>
__global__ __launch_bounds__(64)
void k_sweep_wave(TileView G, SweepData<T> dat, unsigned int* flags,
                  unsigned long long* ticket, unsigned long long ticketBase,
                  unsigned long long epoch, DevState* st, FusedFinalize F)
{
    constexpr int BX = WAVE_BX, BY = WAVE_BX, BZ = WAVE_BZ, NC = BX*BY*BZ;
    // staged inputs, premultiplied: a = rD*x (forward) or w (backward),
    // c[e] = rD*coefficient (the products of the reference expressions
    // wA = rD*rA and wA -= rD*coefficient*wA(nb), formed once here)
    __shared__ T sA[NC];
    __shared__ T sC[3][NC];
    __shared__ T haloK[BX*BY], haloJ[BX*BZ], haloI[BY*BZ];
    const int lane = threadIdx.x;
    const int nTiles = G.nTI*G.nTJ*G.nTK;

    int tile = -1;
    if (lane == 0)
    {
        const unsigned long long t = atomicAdd(ticket, 1ull) - ticketBase;
        if (t < (unsigned long long)nTiles)
        {
            const int ti = (int)t;
            tile = BWD ? G.order[nTiles - 1 - ti] : G.order[ti];
        }
        else if (st)
        {
            st->scheduleError = 1;
        }
    }
    tile = __builtin_amdgcn_readfirstlane(__shfl(tile, 0, 64));   // uniform
    if (tile < 0) return;
    const unsigned int doneVal = (unsigned int)(epoch + 1);
    if (st && st->done)
    {
        if (lane == 0) wv_st_coh(&flags[tile], doneVal);
        return;
    }

    const int TI = tile % G.nTI, TJ = (tile/G.nTI) % G.nTJ, TK = tile/(G.nTI*G.nTJ);
    const int x0 = TI*BX, y0 = TJ*BY, z0 = TK*BZ;
    const int ex = min(BX, G.nx - x0), ey = min(BY, G.ny - y0), ez = min(BZ, G.nz - z0);
    const int64_t sy = G.nx, sz = (int64_t)G.nx*G.ny;
    const int64_t base = x0 + sy*y0 + sz*z0;
    const int il = lane % BX, jl = lane/BX;
    const bool col = il < ex && jl < ey;
    const int gi = x0 + il, gj = y0 + jl;
    const int64_t cBase = base + il + sy*jl;
    const bool wantDot = BWD && dat.dotR != nullptr;

    // ---- stage the pencil (unconditional loads, all in flight) --------------
    T rr[BWD ? BZ : 1];
    {
        T ra[BZ], rd[BZ], rc[3][BZ];
        const bool sh = dat.dirShift != 0;
        const int64_t off[3] = {sh ? sz : 0, sh ? sy : 0, sh ? 1 : 0};
#pragma unroll
        for (int kk = 0; kk < BZ; kk++)
        {
            const bool ok = col && kk < ez;
            const int gk = z0 + kk;
            const int64_t c = ok ? cBase + sz*kk : 0;
            // neighbours on the pass's side: k, j, i
            const bool has[3] = {BWD ? gk < G.nz - 1 : gk > 0,
                                 BWD ? gj < G.ny - 1 : gj > 0,
                                 BWD ? gi < G.nx - 1 : gi > 0};
#pragma unroll
            for (int e = 0; e < 3; e++)
            {
                const bool h = ok && has[e];
                const T v = dat.dir[e][h ? c + off[e] : 0];
                rc[e][kk] = h ? v : T(0);
            }
            ra[kk] = BWD ? dat.w[c] : dat.x[c];
            rd[kk] = dat.rD[c];
            if constexpr (BWD) rr[kk] = wantDot ? dat.dotR[c] : T(0);
        }
#pragma unroll
        for (int kk = 0; kk < BZ; kk++)
        {
            const int q = lane + BX*BY*kk;
            sA[q] = BWD ? ra[kk] : rd[kk]*ra[kk];
            sC[0][q] = rd[kk]*rc[0][kk];
            sC[1][q] = rd[kk]*rc[1][kk];
            sC[2][q] = rd[kk]*rc[2][kk];
        }
    }

    // ---- wait for the upstream tiles ----------------------------------------
    if (lane < 3)
    {
        int nbI = TI, nbJ = TJ, nbK = TK;
        bool has;
        if (lane == 0) { nbI += BWD ? 1 : -1; has = BWD ? nbI < G.nTI : nbI >= 0; }
        else if (lane == 1) { nbJ += BWD ? 1 : -1; has = BWD ? nbJ < G.nTJ : nbJ >= 0; }
        else { nbK += BWD ? 1 : -1; has = BWD ? nbK < G.nTK : nbK >= 0; }
        if (has)
        {
            const int nb = nbI + G.nTI*(nbJ + G.nTJ*nbK);
            uint32_t spins = 0;
            while (wv_ld_coh(&flags[nb]) < doneVal)
            {
                if (++spins > WAVE_SPIN_LIMIT)
                {
                    if (st) st->scheduleError = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();

    // ---- dependency halos: the upstream tiles' boundary values --------------
    {
        // k face: lane (il, jl)
        const int gk = BWD ? z0 + ez : z0 - 1;
        const bool okK = col && gk >= 0 && gk < G.nz;
        T vK[1];
        vK[0] = okK ? wv_ld_coh(&dat.w[base + il + sy*jl + sz*(gk - z0)]) : T(0);
        // j face (i, k) and i face (j, k): BX*BZ = BY*BZ = 128 values, two per lane
        T vJ[2], vI[2];
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            const int q = lane + 64*h;
            const int a = q % BX, kl = q/BX;
            const int gjh = BWD ? y0 + ey : y0 - 1;
            const bool okJ = a < ex && kl < ez && gjh >= 0 && gjh < G.ny;
            vJ[h] = okJ ? wv_ld_coh(&dat.w[base + a + sy*(gjh - y0) + sz*kl]) : T(0);
            const int gih = BWD ? x0 + ex : x0 - 1;
            const bool okI = a < ey && kl < ez && gih >= 0 && gih < G.nx;
            vI[h] = okI ? wv_ld_coh(&dat.w[base + (gih - x0) + sy*a + sz*kl]) : T(0);
        }
        haloK[lane] = vK[0];
#pragma unroll
        for (int h = 0; h < 2; h++)
        {
            haloJ[lane + 64*h] = vJ[h];
            haloI[lane + 64*h] = vI[h];
        }
    }
    __syncthreads();

    // ---- hyperplane walk (lane shuffles, no barrier) -------------------------
    const int nSteps = ex + ey + ez - 2;
    const bool inJ = BWD ? jl < ey - 1 : jl > 0;
    const bool inI = BWD ? il < ex - 1 : il > 0;
    const bool hasJ = BWD ? gj < G.ny - 1 : gj > 0;
    const bool hasI = BWD ? gi < G.nx - 1 : gi > 0;
    const int srcJ = BWD ? (jl < BY - 1 ? lane + BX : lane) : (jl > 0 ? lane - BX : lane);
    auto cellAt = [&](int s) -> int
    {
        return BWD ? (ex - 1 - il) + (ey - 1 - jl) + (ez - 1) - s : s - il - jl;
    };
    auto clampK = [&](int kl) -> int { return kl < 0 ? 0 : (kl >= ez ? ez - 1 : kl); };
    // the staged inputs and halo values of a step are read one step ahead;
    // only the exchange of the previous step's results is on the chain
    int qn = lane + BX*BY*clampK(cellAt(0));
    int kn0 = clampK(cellAt(0));
    T nA = sA[qn], nC0 = sC[0][qn], nC1 = sC[1][qn], nC2 = sC[2][qn];
    T nhJ = haloJ[il + BX*kn0], nhI = haloI[jl + BY*kn0];
    const T hK = haloK[lane];
    T prev = T(0);      // this lane's result of the previous step
    T last = T(0);      // this lane's previous cell along k
    for (int s = 0; s < nSteps; s++)
    {
        const int kl = cellAt(s);
        const bool act = col && kl >= 0 && kl < ez;
        const int kc = clampK(kl);
        const int q = lane + BX*BY*kc;
        const int gk = z0 + kc;
        const bool inK = BWD ? kc < ez - 1 : kc > 0;
        const bool hasK = BWD ? gk < G.nz - 1 : gk > 0;
        const T A = nA, C0 = nC0, C1 = nC1, C2 = nC2, hJ = nhJ, hI = nhI;
        // the previous hyperplane: (j-/+1) by lane permute, (i-/+1) by a
        // row shift (lanes il and il -/+ 1 share their 16-lane row)
        const T sJ = wv_shfl(prev, srcJ);
        const T sI = wv_row_shift<!BWD>(prev);
        {
            const int kn = clampK(cellAt(s + 1));
            qn = lane + BX*BY*kn;
            nA = sA[qn]; nC0 = sC[0][qn]; nC1 = sC[1][qn]; nC2 = sC[2][qn];
            nhJ = haloJ[il + BX*kn];
            nhI = haloI[jl + BY*kn];
        }
        // dependency values in reference order k, j, i (forward: lower
        // neighbours ascending; backward: upper neighbours descending)
        const T vK = inK ? last : hK;
        const T vJ = inJ ? sJ : hJ;
        const T vI = inI ? sI : hI;
        T w = A;
        w = hasK ? w - C0*vK : w;
        w = hasJ ? w - C1*vJ : w;
        w = hasI ? w - C2*vI : w;
        prev = act ? w : T(0);
        if (act)
        {
            last = w;
            sA[q] = w;          // own cell, input already consumed
        }
    }

    // ---- the tile's results: whole rows of 8, then publish completion ---------
    __syncthreads();
    {
        // eight rows (j, k) per store, lane & 7 the cell along i
        const int ci = lane & 7;
#pragma unroll
        for (int it = 0; it < BY*BZ/8; it++)
        {
            const int row = (lane >> 3) + 8*it;
            const int rj = row % BY, rk = row/BY;
            if (ci < ex && rj < ey && rk < ez)
                wv_st_coh(&dat.w[base + ci + sy*rj + sz*rk], sA[ci + BX*row]);
        }
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (lane == 0) wv_st_coh(&flags[tile], doneVal);

    // ---- fused sum w*r of the tile (after publishing: off the chain) --------
    if constexpr (BWD)
    {
        if (wantDot)
        {
            double acc = 0.0;
#pragma unroll
            for (int kk = 0; kk < BZ; kk++)
                if (col && kk < ez) acc += (double)sA[lane + BX*BY*kk]*(double)rr[kk];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            partials_arrive(acc, 0.0, dat.partials, nTiles, tile, F);
        }
    }
}

template<class T>
int32_t launch_sweep_wave(bool backward, const TileView& G, const SweepData<T>& d,
                          TileSync& sync, DevState* st, hipStream_t s, const FusedFinalize& F)
{
    if (G.bx != WAVE_BX || G.by != WAVE_BX || G.bz != WAVE_BZ)
        fail(FH_ERR_INVALID, "tile shape does not match the one-wavefront sweep kernels");
    if (!d.dir[0] || !d.dir[1] || !d.dir[2])
        fail(FH_ERR_INVALID, "internal: the one-wavefront sweep needs packed coefficients");
    if (d.dotR && (!backward || !d.partials))
        fail(FH_ERR_INVALID, "internal: fused sum on a sweep that does not provide it");
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    if (backward)
        hipLaunchKernelGGL((k_sweep_wave<T, true>), dim3(nTiles), dim3(64), 0, s, G, d,
                           sync.flags, sync.ticket, sync.ticketBase, sync.epoch, st, F);
    else
        hipLaunchKernelGGL((k_sweep_wave<T, false>), dim3(nTiles), dim3(64), 0, s, G, d,
                           sync.flags, sync.ticket, sync.ticketBase, sync.epoch, st, F);
    sync.ticketBase += (unsigned long long)nTiles;
    sync.epoch++;
    return nTiles;
}

template int32_t launch_sweep_wave<double>(bool, const TileView&, const SweepData<double>&,
                                           TileSync&, DevState*, hipStream_t,
                                           const FusedFinalize&);
template int32_t launch_sweep_wave<float>(bool, const TileView&, const SweepData<float>&,
                                          TileSync&, DevState*, hipStream_t,
                                           const FusedFinalize&);

} // namespace fh
