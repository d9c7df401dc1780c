// Tiled wavefront sweeps for hex-box meshes (natural blockMesh numbering).
//
// The reference recurrences (DICPreconditioner.C:87-123, DILUPreconditioner.C:
// 88-135, the DIC/DILU factor loops, GaussSeidelSmoother.C:66-187,
// symGaussSeidelSmoother.C:66-216) make cell (i,j,k) wait for its lower
// neighbours (i-1,j,k), (i,j-1,k), (i,j,k-1) (forward) or upper neighbours
// (backward).  The box is cut into BX x BY x BZ tiles; one workgroup owns a
// tile, thread (i,j) owns the tile's pencil along k, and the workgroup walks
// the tile's hyperplanes i+j+k = s with one LDS barrier per step:
//   k-1 neighbour   previous result of the same thread (register)
//   i-1, j-1        LDS step buffer written by neighbouring threads at s-1
//   tile boundary   halo staged from the upstream tile once it is complete
// A tile starts when its (up to three) upstream tiles have published their
// completion flag; tiles are handed out by a ticket counter in wavefront
// order, so a tile only waits for tiles that already run.  Tile outputs and
// halo reads use device-coherent (sc1) accesses, and a tile waits for its
// stores (s_waitcnt vmcnt(0)) before publishing, so no L2 writeback or
// invalidation is needed.  Each row keeps the reference face order, so the
// results are bitwise those of the sequential loops.
#include "fh_tile_sweep.h"

namespace fh
{

#define TILE_SPIN_LIMIT (1u << 20)

template<class T>
__device__ __forceinline__ T ld_coh(const T* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template<class T>
__device__ __forceinline__ void st_coh(T* p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// wait until this thread's global stores have completed
__device__ __forceinline__ void wait_stores()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

__device__ __forceinline__ int32_t sell_at(const int32_t* sliceStart, int64_t c, int32_t k)
{
    return sliceStart[c >> 6] + (int32_t)(c & 63) + (k << 6);
}

// Which kinds read what.  FWD kinds depend on lower neighbours and walk
// hyperplanes upwards; BWD kinds depend on upper neighbours and walk down.
template<int KIND> struct TileKind
{
    static constexpr bool bwd = (KIND == SWEEP_PRECOND_BWD || KIND == SWEEP_GS_BWD);
    static constexpr bool gs = (KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD);
    static constexpr bool rd = (KIND == SWEEP_RD_DIC || KIND == SWEEP_RD_DILU);
};

// LDS layout: staged per-cell inputs of the tile and the halo / step buffers
template<class T, int BX, int BY, int BZ, int KIND>
struct TileShared
{
    static constexpr int NC = BX*BY*BZ;
    T a[NC];                      // x (precond fwd), w (precond bwd), diag (rd), b (gs)
    T d[NC];                      // rD (precond), 1/.. unused (rd), diag (gs)
    T c[3][NC];                   // dependency-side coefficients, k, j, i
    T e[TileKind<KIND>::gs || KIND == SWEEP_RD_DILU ? 3 : 1]
       [TileKind<KIND>::gs || KIND == SWEEP_RD_DILU ? NC : 1];
                                  // gs: other-side coefficients; dilu: partner
    T o[TileKind<KIND>::gs ? NC : 1];          // gs: staged old psi of the tile
    T haloK[BX*BY];               // dependency halos (new values)
    T haloJ[BX*BZ];
    T haloI[BY*BZ];
    T oK[TileKind<KIND>::gs ? BX*BY : 1];      // gs: other-side halos (old values)
    T oJ[TileKind<KIND>::gs ? BX*BZ : 1];
    T oI[TileKind<KIND>::gs ? BY*BZ : 1];
    T step[2][BX*BY];
    int32_t tile;
};

// LOCAL: every tile is factorised and solved on its own (block-Jacobi over
// tiles, couplings between tiles dropped) -- no waits, all tiles in parallel
template<class T, int KIND, int BX, int BY, int BZ, bool LOCAL>
__global__ __launch_bounds__(BX*BY)
void k_sweep_tile(TileView G, RowsView R, SweepData<T> dat, unsigned int* flags,
                  unsigned long long* ticket, unsigned long long ticketBase,
                  unsigned long long epoch, DevState* st)
{
    typedef TileKind<KIND> K;
    __shared__ TileShared<T, BX, BY, BZ, KIND> S;
    const int tid = threadIdx.x;
    const int nTiles = G.nTI*G.nTJ*G.nTK;

    if (LOCAL)
    {
        if (tid == 0) S.tile = blockIdx.x;
    }
    else if (tid == 0)
    {
        const unsigned long long t = atomicAdd(ticket, 1ull) - ticketBase;
        int tile = -1;
        if (t < (unsigned long long)nTiles)
        {
            const int ti = (int)t;
            tile = K::bwd ? G.order[nTiles - 1 - ti] : G.order[ti];
        }
        else if (st)
        {
            st->scheduleError = 1;
        }
        S.tile = tile;
    }
    __syncthreads();
    const int tile = S.tile;
    if (tile < 0) return;
    const unsigned int doneVal = (unsigned int)(epoch + 1);
    if (st && st->done)
    {
        if (!LOCAL && tid == 0) st_coh(&flags[tile], doneVal);
        return;
    }

    const int TI = tile % G.nTI, TJ = (tile/G.nTI) % G.nTJ, TK = tile/(G.nTI*G.nTJ);
    const int x0 = TI*BX, y0 = TJ*BY, z0 = TK*BZ;
    const int ex = min(BX, G.nx - x0), ey = min(BY, G.ny - y0), ez = min(BZ, G.nz - z0);
    const int64_t sy = G.nx, sz = (int64_t)G.nx*G.ny;
    const int64_t base = x0 + sy*y0 + sz*z0;

    // ---- stage the tile's inputs (independent of upstream tiles) ----------
    for (int q = tid; q < BX*BY*BZ; q += BX*BY)
    {
        const int il = q % BX, jl = (q/BX) % BY, kl = q/(BX*BY);
        if (il >= ex || jl >= ey || kl >= ez) continue;
        const int gi = x0 + il, gj = y0 + jl, gk = z0 + kl;
        const int64_t c = base + il + sy*jl + sz*kl;
        const int32_t info = R.rowInfo[c];
        const int32_t nLo = info & 0xffff;
        // lower entries in face order: k-1, j-1, i-1 (those that exist)
        int eK = -1, eJ = -1, eI = -1, n = 0;
        if (gk > 0) eK = n++;
        if (gj > 0) eJ = n++;
        if (gi > 0) eI = n++;
        // upper entries after the lower ones: i+1, j+1, k+1
        int uI = -1, uJ = -1, uK = -1;
        n = nLo;
        if (gi < G.nx - 1) uI = n++;
        if (gj < G.ny - 1) uJ = n++;
        if (gk < G.nz - 1) uK = n++;
        const int dK = K::bwd ? uK : eK, dJ = K::bwd ? uJ : eJ, dI = K::bwd ? uI : eI;
        S.c[0][q] = dK >= 0 ? dat.val[sell_at(R.sliceStart, c, dK)] : T(0);
        S.c[1][q] = dJ >= 0 ? dat.val[sell_at(R.sliceStart, c, dJ)] : T(0);
        S.c[2][q] = dI >= 0 ? dat.val[sell_at(R.sliceStart, c, dI)] : T(0);
        if constexpr (K::gs)
        {
            // other side: upper (forward) in i, j, k order / lower (backward)
            // in k, j, i order
            const int oA = K::bwd ? eK : uI, oB = K::bwd ? eJ : uJ, oC = K::bwd ? eI : uK;
            S.e[0][q] = oA >= 0 ? dat.val[sell_at(R.sliceStart, c, oA)] : T(0);
            S.e[1][q] = oB >= 0 ? dat.val[sell_at(R.sliceStart, c, oB)] : T(0);
            S.e[2][q] = oC >= 0 ? dat.val[sell_at(R.sliceStart, c, oC)] : T(0);
            S.o[q] = dat.w[c];
            S.a[q] = dat.x[c];
            S.d[q] = dat.diag[c];
        }
        else if constexpr (K::rd)
        {
            S.a[q] = dat.diag[c];
            if constexpr (KIND == SWEEP_RD_DILU)
            {
                // partner: upper[f] of the same faces (DILUPreconditioner.C:67-72)
                S.e[0][q] = eK >= 0 ? (T)dat.upperF[R.face[sell_at(R.sliceStart, c, eK)] >> 1] : T(0);
                S.e[1][q] = eJ >= 0 ? (T)dat.upperF[R.face[sell_at(R.sliceStart, c, eJ)] >> 1] : T(0);
                S.e[2][q] = eI >= 0 ? (T)dat.upperF[R.face[sell_at(R.sliceStart, c, eI)] >> 1] : T(0);
            }
        }
        else
        {
            S.a[q] = K::bwd ? dat.w[c] : dat.x[c];
            S.d[q] = dat.rD[c];
        }
    }
    if constexpr (K::gs)
    {
        // other-side halos: cells of downstream (forward) / upstream
        // (backward) tiles, which this tile's completion still guards
        for (int q = tid; q < BX*BY; q += BX*BY)
        {
            const int il = q % BX, jl = q/BX;
            const int gk = K::bwd ? z0 - 1 : z0 + ez;
            if (il < ex && jl < ey && gk >= 0 && gk < G.nz)
                S.oK[q] = dat.w[base + il + sy*jl + sz*(gk - z0)];
        }
        for (int q = tid; q < BX*BZ; q += BX*BY)
        {
            const int il = q % BX, kl = q/BX;
            const int gj = K::bwd ? y0 - 1 : y0 + ey;
            if (il < ex && kl < ez && gj >= 0 && gj < G.ny)
                S.oJ[q] = dat.w[base + il + sy*(gj - y0) + sz*kl];
        }
        for (int q = tid; q < BY*BZ; q += BX*BY)
        {
            const int jl = q % BY, kl = q/BY;
            const int gi = K::bwd ? x0 - 1 : x0 + ex;
            if (jl < ey && kl < ez && gi >= 0 && gi < G.nx)
                S.oI[q] = dat.w[base + (gi - x0) + sy*jl + sz*kl];
        }
    }

    // ---- wait for the upstream tiles ----------------------------------------
    if (!LOCAL && tid < 3)
    {
        int nbI = TI, nbJ = TJ, nbK = TK;
        bool has;
        if (tid == 0) { nbI += K::bwd ? 1 : -1; has = K::bwd ? nbI < G.nTI : nbI >= 0; }
        else if (tid == 1) { nbJ += K::bwd ? 1 : -1; has = K::bwd ? nbJ < G.nTJ : nbJ >= 0; }
        else { nbK += K::bwd ? 1 : -1; has = K::bwd ? nbK < G.nTK : nbK >= 0; }
        if (has)
        {
            const int nb = nbI + G.nTI*(nbJ + G.nTJ*nbK);
            uint32_t spins = 0;
            while (ld_coh(&flags[nb]) < doneVal)
            {
                if (++spins > TILE_SPIN_LIMIT)
                {
                    if (st) st->scheduleError = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
    }
    __syncthreads();

    // ---- dependency halos (values the upstream tiles just produced) --------
    if (!LOCAL)
    {
    for (int q = tid; q < BX*BY; q += BX*BY)
    {
        const int il = q % BX, jl = q/BX;
        const int gk = K::bwd ? z0 + ez : z0 - 1;
        if (il < ex && jl < ey && gk >= 0 && gk < G.nz)
            S.haloK[q] = ld_coh(&dat.w[base + il + sy*jl + sz*(gk - z0)]);
    }
    for (int q = tid; q < BX*BZ; q += BX*BY)
    {
        const int il = q % BX, kl = q/BX;
        const int gj = K::bwd ? y0 + ey : y0 - 1;
        if (il < ex && kl < ez && gj >= 0 && gj < G.ny)
            S.haloJ[q] = ld_coh(&dat.w[base + il + sy*(gj - y0) + sz*kl]);
    }
    for (int q = tid; q < BY*BZ; q += BX*BY)
    {
        const int jl = q % BY, kl = q/BY;
        const int gi = K::bwd ? x0 + ex : x0 - 1;
        if (jl < ey && kl < ez && gi >= 0 && gi < G.nx)
            S.haloI[q] = ld_coh(&dat.w[base + (gi - x0) + sy*jl + sz*kl]);
    }
    }
    __syncthreads();

    // ---- hyperplane walk ------------------------------------------------------
    const int il = tid % BX, jl = tid/BX;
    const int nSteps = ex + ey + ez - 2;
    T prev = T(0);           // result of this pencil's previous cell
    const bool col = il < ex && jl < ey;
    const int gi = x0 + il, gj = y0 + jl;
    for (int s = 0; s < nSteps; s++)
    {
        // forward: s = il + jl + kl; backward walks from the far corner
        const int kl = K::bwd ? (ex - 1 - il) + (ey - 1 - jl) + (ez - 1) - s
                              : s - il - jl;
        const bool act = col && kl >= 0 && kl < ez;
        const T* prevStep = S.step[(s + 1) & 1];
        T out = T(0);
        if (act)
        {
            const int q = il + BX*(jl + BY*kl);
            const int gk = z0 + kl;
            // dependency values in reference order k, j, i (forward: lower
            // neighbours ascending; backward: upper neighbours descending)
            T vK = T(0), vJ = T(0), vI = T(0);
            const bool inK = K::bwd ? kl < ez - 1 : kl > 0;
            const bool inJ = K::bwd ? jl < ey - 1 : jl > 0;
            const bool inI = K::bwd ? il < ex - 1 : il > 0;
            const bool hasK = LOCAL ? inK : (K::bwd ? gk < G.nz - 1 : gk > 0);
            const bool hasJ = LOCAL ? inJ : (K::bwd ? gj < G.ny - 1 : gj > 0);
            const bool hasI = LOCAL ? inI : (K::bwd ? gi < G.nx - 1 : gi > 0);
            if (hasK) vK = inK ? prev : S.haloK[il + BX*jl];
            if (hasJ) vJ = inJ ? prevStep[tid + (K::bwd ? BX : -BX)] : S.haloJ[il + BX*kl];
            if (hasI) vI = inI ? prevStep[tid + (K::bwd ? 1 : -1)] : S.haloI[jl + BY*kl];

            if constexpr (KIND == SWEEP_PRECOND_FWD || KIND == SWEEP_PRECOND_BWD)
            {
                const T rDr = S.d[q];
                T w = K::bwd ? S.a[q] : rDr*S.a[q];
                if (hasK) w -= rDr*S.c[0][q]*vK;
                if (hasJ) w -= rDr*S.c[1][q]*vJ;
                if (hasI) w -= rDr*S.c[2][q]*vI;
                out = w;
            }
            else if constexpr (KIND == SWEEP_RD_DIC)
            {
                T v = S.a[q];
                if (hasK) { const T u = S.c[0][q]; v -= u*u/vK; }
                if (hasJ) { const T u = S.c[1][q]; v -= u*u/vJ; }
                if (hasI) { const T u = S.c[2][q]; v -= u*u/vI; }
                out = v;
            }
            else if constexpr (KIND == SWEEP_RD_DILU)
            {
                T v = S.a[q];
                if (hasK) v -= S.e[0][q]*S.c[0][q]/vK;
                if (hasJ) v -= S.e[1][q]*S.c[1][q]/vJ;
                if (hasI) v -= S.e[2][q]*S.c[2][q]/vI;
                out = v;
            }
            else // Gauss-Seidel
            {
                // old (forward: upper) / forward-pass (backward: lower) values
                const bool oHasI = K::bwd ? gi > 0 : gi < G.nx - 1;
                const bool oHasJ = K::bwd ? gj > 0 : gj < G.ny - 1;
                const bool oHasK = K::bwd ? gk > 0 : gk < G.nz - 1;
                const bool oInI = K::bwd ? il > 0 : il < ex - 1;
                const bool oInJ = K::bwd ? jl > 0 : jl < ey - 1;
                const bool oInK = K::bwd ? kl > 0 : kl < ez - 1;
                T oVI = T(0), oVJ = T(0), oVK = T(0);
                if (oHasI) oVI = oInI ? S.o[q + (K::bwd ? -1 : 1)] : S.oI[jl + BY*kl];
                if (oHasJ) oVJ = oInJ ? S.o[q + (K::bwd ? -BX : BX)] : S.oJ[il + BX*kl];
                if (oHasK) oVK = oInK ? S.o[q + (K::bwd ? -BX*BY : BX*BY)] : S.oK[il + BX*jl];
                T t = S.a[q];
                if (!K::bwd)
                {
                    // bPrime: lower products k, j, i (new values); then the
                    // upper products i, j, k (old values)
                    if (hasK) t -= S.c[0][q]*vK;
                    if (hasJ) t -= S.c[1][q]*vJ;
                    if (hasI) t -= S.c[2][q]*vI;
                    if (oHasI) t -= S.e[0][q]*oVI;
                    if (oHasJ) t -= S.e[1][q]*oVJ;
                    if (oHasK) t -= S.e[2][q]*oVK;
                }
                else
                {
                    // bPrime from the forward pass: lower products k, j, i
                    // (forward values); then upper products i, j, k (new)
                    if (oHasK) t -= S.e[0][q]*oVK;
                    if (oHasJ) t -= S.e[1][q]*oVJ;
                    if (oHasI) t -= S.e[2][q]*oVI;
                    if (hasI) t -= S.c[2][q]*vI;
                    if (hasJ) t -= S.c[1][q]*vJ;
                    if (hasK) t -= S.c[0][q]*vK;
                }
                out = t/S.d[q];
            }
            prev = out;
            st_coh(&dat.w[base + il + sy*jl + sz*kl], out);
        }
        S.step[s & 1][tid] = out;
        __syncthreads();
    }

    // ---- publish completion -------------------------------------------------
    if (!LOCAL)
    {
        wait_stores();
        __syncthreads();
        if (tid == 0) st_coh(&flags[tile], doneVal);
    }
}

// ---------------------------------------------------------------------------

template<class T, int KIND, int BX, int BY, int BZ>
static void launch_tile_kind(const TileView& G, const RowsView& R, const SweepData<T>& d,
                             TileSync& sync, DevState* st, hipStream_t s, bool local)
{
    const int nTiles = G.nTI*G.nTJ*G.nTK;
    if (local)
    {
        if constexpr (!TileKind<KIND>::gs)
            hipLaunchKernelGGL((k_sweep_tile<T, KIND, BX, BY, BZ, true>), dim3(nTiles),
                               dim3(BX*BY), 0, s, G, R, d, sync.flags, sync.ticket,
                               sync.ticketBase, sync.epoch, st);
        else
            fail(FH_ERR_INVALID, "tile-local Gauss-Seidel is not provided");
        return;
    }
    hipLaunchKernelGGL((k_sweep_tile<T, KIND, BX, BY, BZ, false>), dim3(nTiles), dim3(BX*BY),
                       0, s, G, R, d, sync.flags, sync.ticket, sync.ticketBase, sync.epoch,
                       st);
    sync.ticketBase += (unsigned long long)nTiles;
    sync.epoch++;
}

template<class T>
void launch_tile_sweep(int kind, const TileView& G, const RowsView& R,
                       const SweepData<T>& d, TileSync& sync, DevState* st, hipStream_t s,
                       bool local)
{
    if (G.bx != 16 || G.by != 16 || G.bz != tile_bz<T>(kind))
        fail(FH_ERR_INVALID, "tile shape does not match the compiled sweep kernels");
    // on a box the faces of an owner are in ascending neighbour order, so
    // the reversed-losort walk of preconditionT is the reversed face walk
    if (kind == SWEEP_PRECOND_BWD_T) kind = SWEEP_PRECOND_BWD;
    switch (kind)
    {
        case SWEEP_PRECOND_FWD: launch_tile_kind<T, SWEEP_PRECOND_FWD, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local); break;
        case SWEEP_PRECOND_BWD: launch_tile_kind<T, SWEEP_PRECOND_BWD, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local); break;
        case SWEEP_RD_DIC:      launch_tile_kind<T, SWEEP_RD_DIC, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local); break;
        case SWEEP_RD_DILU:     launch_tile_kind<T, SWEEP_RD_DILU, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local); break;
        case SWEEP_GS_FWD:      launch_tile_kind<T, SWEEP_GS_FWD, 16, 16, TileBZ<T>::gs>(G, R, d, sync, st, s, local); break;
        case SWEEP_GS_BWD:      launch_tile_kind<T, SWEEP_GS_BWD, 16, 16, TileBZ<T>::gs>(G, R, d, sync, st, s, local); break;
        default: fail(FH_ERR_INVALID, "unknown sweep kind %d", kind);
    }
}

template void launch_tile_sweep<double>(int, const TileView&, const RowsView&, const SweepData<double>&, TileSync&, DevState*, hipStream_t, bool);
template void launch_tile_sweep<float>(int, const TileView&, const RowsView&, const SweepData<float>&, TileSync&, DevState*, hipStream_t, bool);

} // namespace fh
