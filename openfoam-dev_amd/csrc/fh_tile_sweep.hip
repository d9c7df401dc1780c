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
    // precond: a = rD*x (forward) or w (backward), c = rD*coefficient
    T a[NC];                      // precond (above), diag (rd), b (gs)
    T d[TileKind<KIND>::gs ? NC : 1];          // gs: diag
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
    // Thread (il, jl) stages its own pencil, cells kl = 0..ez-1.  The loads
    // are unconditional (a cell outside the tile or an absent neighbour
    // reads index 0 and its value is replaced by 0), so every load of a
    // phase is in flight before the first is used.  On a box the entry
    // slots of a row follow from its position: the lower entries k-1, j-1,
    // i-1 (those that exist) in face order, then the upper entries i+1,
    // j+1, k+1.
    const int il = tid % BX, jl = tid/BX;
    const bool col = il < ex && jl < ey;
    const int gi = x0 + il, gj = y0 + jl;
    const int64_t cBase = base + il + sy*jl;
    constexpr bool PRE = KIND == SWEEP_PRECOND_FWD || KIND == SWEEP_PRECOND_BWD;
    const bool useDir = PRE && dat.dir[0] != nullptr;
    const bool wantDot = KIND == SWEEP_PRECOND_BWD && dat.dotR != nullptr;
    T rr[KIND == SWEEP_PRECOND_BWD ? BZ : 1];   // dotR of the pencil
    {
        int32_t sb[BZ] = {};
        if (!useDir)
        {
#pragma unroll
            for (int kk = 0; kk < BZ; kk++)
            {
                const int64_t c = (col && kk < ez) ? cBase + sz*kk : 0;
                sb[kk] = R.sliceStart[c >> 6] + (int32_t)(c & 63);
            }
        }
        T ra[BZ], rd[BZ], rc[3][BZ];
        T re[K::gs || KIND == SWEEP_RD_DILU ? 3 : 1][BZ];
        T ro[K::gs ? BZ : 1];
        int32_t fc[KIND == SWEEP_RD_DILU ? 3 : 1][BZ];
#pragma unroll
        for (int kk = 0; kk < BZ; kk++)
        {
            const bool ok = col && kk < ez;
            const int gk = z0 + kk;
            const int64_t c = ok ? cBase + sz*kk : 0;
            const int nLo = (gk > 0) + (gj > 0) + (gi > 0);
            const int eK = gk > 0 ? 0 : -1;
            const int eJ = gj > 0 ? (gk > 0) : -1;
            const int eI = gi > 0 ? (gk > 0) + (gj > 0) : -1;
            const int uI = gi < G.nx - 1 ? nLo : -1;
            const int uJ = gj < G.ny - 1 ? nLo + (gi < G.nx - 1) : -1;
            const int uK = gk < G.nz - 1 ? nLo + (gi < G.nx - 1) + (gj < G.ny - 1) : -1;
            const int dK = K::bwd ? uK : eK, dJ = K::bwd ? uJ : eJ, dI = K::bwd ? uI : eI;
            // SELL value of entry slot e of this row (0 when absent)
            auto sell = [&](int e) -> T
            {
                const bool h = ok && e >= 0;
                const T v = dat.val[h ? sb[kk] + (e << 6) : 0];
                return h ? v : T(0);
            };
            {
                // packed by direction (at the cell, or at the neighbour for
                // dirShift) or the row's SELL entries: one pointer and index
                // per coefficient, no branch
                const bool sh = dat.dirShift != 0;
                const int dd[3] = {dK, dJ, dI};
                const int64_t off[3] = {sh ? sz : 0, sh ? sy : 0, sh ? 1 : 0};
#pragma unroll
                for (int e = 0; e < 3; e++)
                {
                    const bool h = ok && dd[e] >= 0;
                    const T* src = useDir ? dat.dir[e] : dat.val;
                    const int64_t ix = useDir ? c + off[e] : (int64_t)sb[kk] + (dd[e] << 6);
                    const T v = src[h ? ix : 0];
                    rc[e][kk] = h ? v : T(0);
                }
            }
            if constexpr (K::gs)
            {
                // other side: upper (forward) in i, j, k order / lower
                // (backward) in k, j, i order
                re[0][kk] = sell(K::bwd ? eK : uI);
                re[1][kk] = sell(K::bwd ? eJ : uJ);
                re[2][kk] = sell(K::bwd ? eI : uK);
                ro[kk] = dat.w[c];
                ra[kk] = dat.x[c];
                rd[kk] = dat.diag[c];
            }
            else if constexpr (K::rd)
            {
                ra[kk] = dat.diag[c];
                if constexpr (KIND == SWEEP_RD_DILU)
                {
                    const int es[3] = {eK, eJ, eI};
#pragma unroll
                    for (int e = 0; e < 3; e++)
                    {
                        const bool h = ok && es[e] >= 0;
                        const int32_t f = R.face[h ? sb[kk] + (es[e] << 6) : 0] >> 1;
                        fc[e][kk] = h ? f : -1;
                    }
                }
            }
            else
            {
                ra[kk] = K::bwd ? dat.w[c] : dat.x[c];
                rd[kk] = dat.rD[c];
                if constexpr (KIND == SWEEP_PRECOND_BWD)
                    rr[kk] = wantDot ? dat.dotR[c] : T(0);
            }
        }
        if constexpr (KIND == SWEEP_RD_DILU)
        {
            // partner: upper[f] of the same faces (DILUPreconditioner.C:67-72)
#pragma unroll
            for (int kk = 0; kk < BZ; kk++)
#pragma unroll
                for (int e = 0; e < 3; e++)
                {
                    const T v = (T)dat.upperF[fc[e][kk] >= 0 ? fc[e][kk] : 0];
                    re[e][kk] = fc[e][kk] >= 0 ? v : T(0);
                }
        }
#pragma unroll
        for (int kk = 0; kk < BZ; kk++)
        {
            const int q = tid + BX*BY*kk;
            if constexpr (PRE)
            {
                // the products rD*x and rD*coefficient of the reference
                // expressions wA = rD*rA and wA -= rD*coefficient*wA(nb),
                // formed once here (the same values)
                S.a[q] = K::bwd ? ra[kk] : rd[kk]*ra[kk];
                S.c[0][q] = rd[kk]*rc[0][kk];
                S.c[1][q] = rd[kk]*rc[1][kk];
                S.c[2][q] = rd[kk]*rc[2][kk];
            }
            else
            {
                S.a[q] = ra[kk];
                if constexpr (K::gs) S.d[q] = rd[kk];
                S.c[0][q] = rc[0][kk];
                S.c[1][q] = rc[1][kk];
                S.c[2][q] = rc[2][kk];
            }
            if constexpr (K::gs || KIND == SWEEP_RD_DILU)
            {
                S.e[0][q] = re[0][kk];
                S.e[1][q] = re[1][kk];
                S.e[2][q] = re[2][kk];
            }
            if constexpr (K::gs) S.o[q] = ro[kk];
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
    // Per step one batch of LDS reads (the neighbours of the previous step
    // and the halos; clamped indices, values selected afterwards), then the
    // row update in the reference order.  The preconditioner kinds read the
    // next cell's staged inputs one step ahead.
    const int nSteps = ex + ey + ez - 2;
    T prev = T(0);           // result of this pencil's previous cell
    // neighbour slots in the step buffer (own slot where there is none)
    const int tJ = K::bwd ? (jl < BY - 1 ? tid + BX : tid) : (jl > 0 ? tid - BX : tid);
    const int tI = K::bwd ? (il < BX - 1 ? tid + 1 : tid) : (il > 0 ? tid - 1 : tid);
    const bool inJ = K::bwd ? jl < ey - 1 : jl > 0;
    const bool inI = K::bwd ? il < ex - 1 : il > 0;
    const bool hasJ = LOCAL ? inJ : (K::bwd ? gj < G.ny - 1 : gj > 0);
    const bool hasI = LOCAL ? inI : (K::bwd ? gi < G.nx - 1 : gi > 0);
    // the pencil's cell at step s (clamped into the tile)
    auto cellAt = [&](int s) -> int
    {
        const int kl = K::bwd ? (ex - 1 - il) + (ey - 1 - jl) + (ez - 1) - s : s - il - jl;
        return kl;
    };
    if constexpr (PRE)
    {
        int kc0 = cellAt(0);
        kc0 = kc0 < 0 ? 0 : (kc0 >= ez ? ez - 1 : kc0);
        int qn = il + BX*(jl + BY*kc0);
        T nA = S.a[qn], nC0 = S.c[0][qn], nC1 = S.c[1][qn], nC2 = S.c[2][qn];
        for (int s = 0; s < nSteps; s++)
        {
            const int kl = cellAt(s);
            const bool act = col && kl >= 0 && kl < ez;
            const int kc = kl < 0 ? 0 : (kl >= ez ? ez - 1 : kl);
            const int q = il + BX*(jl + BY*kc);
            const int gk = z0 + kc;
            const bool inK = K::bwd ? kc < ez - 1 : kc > 0;
            const bool hasK = LOCAL ? inK : (K::bwd ? gk < G.nz - 1 : gk > 0);
            const T* prevStep = S.step[(s + 1) & 1];
            const T A = nA, C0 = nC0, C1 = nC1, C2 = nC2;
            const T sJ = prevStep[tJ], sI = prevStep[tI];
            T hK = T(0), hJ = T(0), hI = T(0);
            if (!LOCAL)
            {
                hK = S.haloK[il + BX*jl];
                hJ = S.haloJ[il + BX*kc];
                hI = S.haloI[jl + BY*kc];
            }
            {
                // the next step's staged inputs
                int kn = cellAt(s + 1);
                kn = kn < 0 ? 0 : (kn >= ez ? ez - 1 : kn);
                qn = il + BX*(jl + BY*kn);
                nA = S.a[qn]; nC0 = S.c[0][qn]; nC1 = S.c[1][qn]; nC2 = S.c[2][qn];
            }
            // dependency values in reference order k, j, i (forward: lower
            // neighbours ascending; backward: upper neighbours descending)
            const T vK = inK ? prev : hK;
            const T vJ = inJ ? sJ : hJ;
            const T vI = inI ? sI : hI;
            T w = A;
            w = hasK ? w - C0*vK : w;
            w = hasJ ? w - C1*vJ : w;
            w = hasI ? w - C2*vI : w;
            const T out = act ? w : T(0);
            if (act)
            {
                prev = out;
                st_coh(&dat.w[base + il + sy*jl + sz*kl], out);
                // the final value, for the fused w.r sum (own cell only)
                if (KIND == SWEEP_PRECOND_BWD && wantDot) S.a[q] = out;
            }
            S.step[s & 1][tid] = out;
            __syncthreads();
        }
    }
    else
    {
    for (int s = 0; s < nSteps; s++)
    {
        const int kl = cellAt(s);
        const bool act = col && kl >= 0 && kl < ez;
        const int kc = kl < 0 ? 0 : (kl >= ez ? ez - 1 : kl);
        const int q = il + BX*(jl + BY*kc);
        const int gk = z0 + kc;
        const T* prevStep = S.step[(s + 1) & 1];
        const bool inK = K::bwd ? kc < ez - 1 : kc > 0;
        const bool hasK = LOCAL ? inK : (K::bwd ? gk < G.nz - 1 : gk > 0);
        // dependency values in reference order k, j, i (forward: lower
        // neighbours ascending; backward: upper neighbours descending)
        const T A = S.a[q];
        const T C0 = S.c[0][q], C1 = S.c[1][q], C2 = S.c[2][q];
        const T sJ = prevStep[tJ], sI = prevStep[tI];
        T hK = T(0), hJ = T(0), hI = T(0);
        if (!LOCAL)
        {
            hK = S.haloK[il + BX*jl];
            hJ = S.haloJ[il + BX*kc];
            hI = S.haloI[jl + BY*kc];
        }
        const T vK = inK ? prev : hK;
        const T vJ = inJ ? sJ : hJ;
        const T vI = inI ? sI : hI;
        T out;
        if constexpr (KIND == SWEEP_RD_DIC)
        {
            T v = A;
            v = hasK ? v - C0*C0/vK : v;
            v = hasJ ? v - C1*C1/vJ : v;
            v = hasI ? v - C2*C2/vI : v;
            out = v;
        }
        else if constexpr (KIND == SWEEP_RD_DILU)
        {
            const T E0 = S.e[0][q], E1 = S.e[1][q], E2 = S.e[2][q];
            T v = A;
            v = hasK ? v - E0*C0/vK : v;
            v = hasJ ? v - E1*C1/vJ : v;
            v = hasI ? v - E2*C2/vI : v;
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
            const bool oInK = K::bwd ? kc > 0 : kc < ez - 1;
            const T E0 = S.e[0][q], E1 = S.e[1][q], E2 = S.e[2][q];
            const T sOI = S.o[oInI ? q + (K::bwd ? -1 : 1) : q];
            const T sOJ = S.o[oInJ ? q + (K::bwd ? -BX : BX) : q];
            const T sOK = S.o[oInK ? q + (K::bwd ? -BX*BY : BX*BY) : q];
            const T hOI = S.oI[jl + BY*kc], hOJ = S.oJ[il + BX*kc], hOK = S.oK[il + BX*jl];
            const T oVI = oInI ? sOI : hOI;
            const T oVJ = oInJ ? sOJ : hOJ;
            const T oVK = oInK ? sOK : hOK;
            T t = A;
            if (!K::bwd)
            {
                // bPrime: lower products k, j, i (new values); then the
                // upper products i, j, k (old values)
                t = hasK ? t - C0*vK : t;
                t = hasJ ? t - C1*vJ : t;
                t = hasI ? t - C2*vI : t;
                t = oHasI ? t - E0*oVI : t;
                t = oHasJ ? t - E1*oVJ : t;
                t = oHasK ? t - E2*oVK : t;
            }
            else
            {
                // bPrime from the forward pass: lower products k, j, i
                // (forward values); then upper products i, j, k (new)
                t = oHasK ? t - E0*oVK : t;
                t = oHasJ ? t - E1*oVJ : t;
                t = oHasI ? t - E2*oVI : t;
                t = hasI ? t - C2*vI : t;
                t = hasJ ? t - C1*vJ : t;
                t = hasK ? t - C0*vK : t;
            }
            out = t/S.d[q];
        }
        if (!act) out = T(0);
        if (act)
        {
            prev = out;
            st_coh(&dat.w[base + il + sy*jl + sz*kl], out);
        }
        S.step[s & 1][tid] = out;
        __syncthreads();
    }
    }

    // ---- publish completion -------------------------------------------------
    if (!LOCAL)
    {
        wait_stores();
        __syncthreads();
        if (tid == 0) st_coh(&flags[tile], doneVal);
    }

    // ---- fused sum w*r of the tile (after publishing: off the chain) --------
    if constexpr (KIND == SWEEP_PRECOND_BWD)
    {
        if (wantDot)
        {
            __shared__ double red[BX*BY/64];
            double acc = 0.0;
#pragma unroll
            for (int kk = 0; kk < BZ; kk++)
                if (col && kk < ez) acc += (double)S.a[tid + BX*BY*kk]*(double)rr[kk];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
            if ((tid & 63) == 0) red[tid >> 6] = acc;
            __syncthreads();
            if (tid == 0)
            {
                double t = 0.0;
#pragma unroll
                for (int w = 0; w < BX*BY/64; w++) t += red[w];
                dat.partials[tile] = t;
                dat.partials[nTiles + tile] = 0.0;
            }
        }
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
    // on a box the faces of an owner are in ascending neighbour order, so
    // the reversed-losort walk of preconditionT is the reversed face walk
    if (kind == SWEEP_PRECOND_BWD_T) kind = SWEEP_PRECOND_BWD;
    const bool pre = kind == SWEEP_PRECOND_FWD || kind == SWEEP_PRECOND_BWD;
    // the global preconditioner sweeps also come in the shallower tiles
    // (more tiles resident per CU); tile-local factors fix the shape
    const bool alt = pre && !local && G.bz == TileBZ<T>::alt;
    if (G.bx != 16 || G.by != 16 || !(G.bz == tile_bz<T>(kind) || alt))
        fail(FH_ERR_INVALID, "tile shape does not match the compiled sweep kernels");
    if (d.dotR && (kind != SWEEP_PRECOND_BWD || !d.partials))
        fail(FH_ERR_INVALID, "internal: fused sum on a sweep that does not provide it");
    switch (kind)
    {
        case SWEEP_PRECOND_FWD:
            if (alt) launch_tile_kind<T, SWEEP_PRECOND_FWD, 16, 16, TileBZ<T>::alt>(G, R, d, sync, st, s, false);
            else launch_tile_kind<T, SWEEP_PRECOND_FWD, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local);
            break;
        case SWEEP_PRECOND_BWD:
            if (alt) launch_tile_kind<T, SWEEP_PRECOND_BWD, 16, 16, TileBZ<T>::alt>(G, R, d, sync, st, s, false);
            else launch_tile_kind<T, SWEEP_PRECOND_BWD, 16, 16, TileBZ<T>::pre>(G, R, d, sync, st, s, local);
            break;
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
