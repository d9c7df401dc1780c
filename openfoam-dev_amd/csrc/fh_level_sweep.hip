// One-workgroup pipelined level sweep for one-part meshes (GAMG coarse
// levels and other meshes without a tile structure).
//
// The library numbering of such a mesh is sorted by dependency level, so a
// level is a contiguous range of rows whose reads in the sweep direction lie
// in earlier levels.  One workgroup walks the levels in order with one
// barrier per level.  The latency a level pays is kept off the global
// memory path:
//
//  * the values of the level just finished are kept in LDS (two buffers by
//    level parity) as well as stored to the vector, so the next level reads
//    its most recent dependencies from LDS;
//  * everything else a row reads -- its row data, its coefficients, the
//    values of older levels (final since an earlier barrier) and, for
//    Gauss-Seidel, the not yet updated neighbours of the other direction --
//    is loaded into registers one level ahead, while the workgroup finishes
//    the current level.
//
// A row's arithmetic is that of the sweep's row update (sweep_row): the same
// entries in the same order with the same operations, so the results are
// bitwise those of the chunk and part sweeps.  Rows beyond the first
// blockDim of a level and entries beyond MAXE of a row are read at compute
// time instead (correct, without the prefetch).
#include "fh_kernels.h"

namespace fh
{

namespace
{

constexpr int LVL_THREADS = 1024;

template<class T> struct LvlCfg;
// values of the previous level kept in LDS (two buffers: 2 x 48 KB) and the
// entries a thread prefetches for its row
template<> struct LvlCfg<double> { static constexpr int CAP = 6144; static constexpr int MAXE = 12; };
template<> struct LvlCfg<float> { static constexpr int CAP = 12288; static constexpr int MAXE = 16; };

__device__ __forceinline__ int32_t lvl_base(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

template<int KIND> struct LvlKind;
template<> struct LvlKind<SWEEP_GS_FWD>      { static constexpr bool bwd = false; };
template<> struct LvlKind<SWEEP_GS_BWD>      { static constexpr bool bwd = true; };
template<> struct LvlKind<SWEEP_PRECOND_FWD> { static constexpr bool bwd = false; };
template<> struct LvlKind<SWEEP_PRECOND_BWD> { static constexpr bool bwd = true; };

// the row a thread holds for a level: its scalars and, per used entry in
// the order the update takes them (position p), the coefficient and where
// the value comes from: loaded ahead (slot -1, value in v), the LDS copy of
// the previous level (slot >= 0), or the vector at compute time (slot <= -2:
// column -slot - 2, a previous-level row past the LDS capacity)
template<class T, int MAXE>
struct LvlRow
{
    int32_t r, cnt, base, nLo, nEnt;
    T a, dg;
    T val[MAXE];
    T v[MAXE];
    int32_t slot[MAXE];
};

// entries an update takes, in its order: Gauss-Seidel all entries
// ascending, the forward DIC/DILU pass the lower entries ascending, the
// backward pass the upper entries descending
template<int KIND>
__device__ __forceinline__ int32_t lvl_count(int32_t nLo, int32_t nEnt)
{
    if constexpr (KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD) return nEnt;
    else if constexpr (KIND == SWEEP_PRECOND_FWD) return nLo;
    else return nEnt - nLo;
}
template<int KIND>
__device__ __forceinline__ int32_t lvl_entry(int32_t p, int32_t nEnt)
{
    if constexpr (KIND == SWEEP_PRECOND_BWD) return nEnt - 1 - p;
    else return p;
}

// value of column c read at compute time: the previous level from LDS
// (or the vector past the LDS capacity), anything else from the vector
template<class T, int CAP>
__device__ __forceinline__ T lvl_value(int32_t c, int32_t pLo, int32_t pHi, const T* lds,
                                       const T* w)
{
    if (c >= pLo && c < pHi && c - pLo < CAP) return lds[c - pLo];
    return w[c];
}

// row r of a level whose previous level (in sweep order) is rows
// [pLo, pHi), still being computed: load everything else
template<class T, int KIND, int MAXE, int CAP>
__device__ __forceinline__ void lvl_fetch(const RowsView& R, const SweepData<T>& d, int32_t r,
                                          int32_t pLo, int32_t pHi, LvlRow<T, MAXE>& o)
{
    const int32_t info = R.rowInfo[r];
    o.r = r;
    o.nLo = info & 0xffff;
    o.nEnt = info >> 16;
    o.cnt = lvl_count<KIND>(o.nLo, o.nEnt);
    o.base = lvl_base(R.sliceStart, r);
    if constexpr (KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD)
    {
        o.a = d.x[r];
        o.dg = d.diag[r];
    }
    else if constexpr (KIND == SWEEP_PRECOND_FWD)
    {
        o.a = d.x[r];
        o.dg = d.rD[r];
    }
    else
    {
        o.a = d.w[r];          // the forward pass's value of the row
        o.dg = d.rD[r];
    }
#pragma unroll
    for (int p = 0; p < MAXE; p++)
    {
        o.slot[p] = -1;
        o.val[p] = T(0);
        o.v[p] = T(0);
        if (p < o.cnt)
        {
            const int32_t e = o.base + (lvl_entry<KIND>(p, o.nEnt) << 6);
            const int32_t c = R.col[e];
            o.val[p] = d.val[e];
            if (c >= pLo && c < pHi) o.slot[p] = c - pLo < CAP ? c - pLo : -2 - c;
            else o.v[p] = d.w[c];
        }
    }
}

// the row update of sweep_row on a held row; returns the new value
template<class T, int KIND, int MAXE, int CAP>
__device__ __forceinline__ T lvl_compute(const RowsView& R, const SweepData<T>& d,
                                         const LvlRow<T, MAXE>& o, int32_t pLo, int32_t pHi,
                                         const T* lds)
{
    constexpr bool GS = KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD;
    // GaussSeidelSmoother.C:150-176 / symGaussSeidelSmoother.C:151-205:
    // t = b - sum val*psi, then / diag; DICPreconditioner.C:108-122 /
    // DILUPreconditioner.C:117-134: w = rD*r (forward) or w (backward),
    // then w -= rD*coefficient*w(neighbour)
    const T rDr = o.dg;
    T t = GS || KIND == SWEEP_PRECOND_BWD ? o.a : rDr*o.a;
#pragma unroll
    for (int p = 0; p < MAXE; p++)
    {
        if (p < o.cnt)
        {
            const int32_t sl = o.slot[p];
            const T value = sl >= 0 ? lds[sl] : (sl == -1 ? o.v[p] : d.w[-2 - sl]);
            if constexpr (GS) t -= o.val[p]*value;
            else t -= rDr*o.val[p]*value;
        }
    }
    for (int32_t p = MAXE; p < o.cnt; p++)
    {
        const int32_t e = o.base + (lvl_entry<KIND>(p, o.nEnt) << 6);
        const T value = lvl_value<T, CAP>(R.col[e], pLo, pHi, lds, d.w);
        if constexpr (GS) t -= d.val[e]*value;
        else t -= rDr*d.val[e]*value;
    }
    if constexpr (GS) return t/o.dg;
    else return t;
}

// the row update of a row read at compute time (rows past the first
// blockDim of a level)
template<class T, int KIND, int CAP>
__device__ __forceinline__ T lvl_direct(const RowsView& R, const SweepData<T>& d, int32_t r,
                                     int32_t pLo, int32_t pHi, const T* lds)
{
    constexpr bool GS = KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD;
    const int32_t info = R.rowInfo[r];
    const int32_t nLo = info & 0xffff, nEnt = info >> 16;
    const int32_t cnt = lvl_count<KIND>(nLo, nEnt);
    const int32_t base = lvl_base(R.sliceStart, r);
    const T rDr = GS ? d.diag[r] : d.rD[r];
    T t = GS ? d.x[r] : (KIND == SWEEP_PRECOND_BWD ? d.w[r] : rDr*d.x[r]);
    for (int32_t p = 0; p < cnt; p++)
    {
        const int32_t e = base + (lvl_entry<KIND>(p, nEnt) << 6);
        const T value = lvl_value<T, CAP>(R.col[e], pLo, pHi, lds, d.w);
        if constexpr (GS) t -= d.val[e]*value;
        else t -= rDr*d.val[e]*value;
    }
    if constexpr (GS) return t/rDr;
    else return t;
}

template<class T, int KIND>
__global__ __launch_bounds__(LVL_THREADS)
void k_sweep_level_wg(RowsView R, SchedView S, SweepData<T> d, DevState* st)
{
    constexpr int CAP = LvlCfg<T>::CAP, MAXE = LvlCfg<T>::MAXE;
    constexpr bool BWD = LvlKind<KIND>::bwd;
    __shared__ T lv[2][CAP];
    if (st && st->done) return;
    const int32_t nSeg = S.partLevelOff[1] - S.partLevelOff[0] - 1;
    const int32_t* seg = S.segStart + S.partLevelOff[0];
    const int32_t tid = (int32_t)threadIdx.x;
    const int32_t nt = (int32_t)blockDim.x;
    // level s in sweep order is library level L(s); its previous level in
    // sweep order supplies the LDS values
    auto levelOf = [&](int32_t s) { return BWD ? nSeg - 1 - s : s; };
    LvlRow<T, MAXE> cur;
    cur.cnt = 0;
    {
        const int32_t L = levelOf(0);
        const int32_t r = seg[L] + tid;
        if (r < seg[L + 1]) lvl_fetch<T, KIND, MAXE, CAP>(R, d, r, 0, 0, cur);
    }
    for (int32_t s = 0; s < nSeg; s++)
    {
        const int32_t L = levelOf(s);
        const int32_t r0 = seg[L], r1 = seg[L + 1];
        // previous level in sweep order (none for the first)
        int32_t pLo = 0, pHi = 0;
        if (s > 0)
        {
            const int32_t P = levelOf(s - 1);
            pLo = seg[P];
            pHi = seg[P + 1];
        }
        const T* ldsPrev = lv[(s + 1) & 1];
        T* ldsMine = lv[s & 1];
        if (r0 + tid < r1)
        {
            const T v = lvl_compute<T, KIND, MAXE, CAP>(R, d, cur, pLo, pHi, ldsPrev);
            d.w[cur.r] = v;
            if (cur.r - r0 < CAP) ldsMine[cur.r - r0] = v;
        }
        for (int32_t r = r0 + tid + nt; r < r1; r += nt)
        {
            const T v = lvl_direct<T, KIND, CAP>(R, d, r, pLo, pHi, ldsPrev);
            d.w[r] = v;
            if (r - r0 < CAP) ldsMine[r - r0] = v;
        }
        // the next level's row, loaded while the workgroup finishes this one
        // (its reads of this level are taken from LDS after the barrier)
        if (s + 1 < nSeg)
        {
            const int32_t N = levelOf(s + 1);
            const int32_t r = seg[N] + tid;
            cur.cnt = 0;
            if (r < seg[N + 1]) lvl_fetch<T, KIND, MAXE, CAP>(R, d, r, r0, r1, cur);
        }
        __syncthreads();
    }
}

template<class T, int KIND>
void launch_kind(const RowsView& R, const SchedView& S, const SweepData<T>& d, DevState* st,
                 hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep_level_wg<T, KIND>), dim3(1), dim3(LVL_THREADS), 0, s, R, S, d,
                       st);
}

} // namespace

template<class T>
bool launch_sweep_level_wg(int kind, const RowsView& R, const SchedView& S,
                           const SweepData<T>& d, DevState* st, hipStream_t s)
{
    if (S.nParts != 1) return false;
    switch (kind)
    {
        case SWEEP_GS_FWD: launch_kind<T, SWEEP_GS_FWD>(R, S, d, st, s); return true;
        case SWEEP_GS_BWD: launch_kind<T, SWEEP_GS_BWD>(R, S, d, st, s); return true;
        case SWEEP_PRECOND_FWD: launch_kind<T, SWEEP_PRECOND_FWD>(R, S, d, st, s); return true;
        case SWEEP_PRECOND_BWD: launch_kind<T, SWEEP_PRECOND_BWD>(R, S, d, st, s); return true;
        default: return false;
    }
}

template bool launch_sweep_level_wg<double>(int, const RowsView&, const SchedView&,
                                            const SweepData<double>&, DevState*, hipStream_t);
template bool launch_sweep_level_wg<float>(int, const RowsView&, const SchedView&,
                                           const SweepData<float>&, DevState*, hipStream_t);

} // namespace fh
