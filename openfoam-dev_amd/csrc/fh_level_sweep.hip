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

constexpr int LVL_THREADS = 512;

template<class T> struct LvlCfg;
// values of a level kept in LDS (two buffers by level parity: 2 x 48 KB) and
// the entries a thread holds in registers for each of its pipelined rows
template<> struct LvlCfg<double> { static constexpr int CAP = 6144; static constexpr int MAXE = 10; };
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
// the value comes from: loaded ahead (slot -1, value in v) or the LDS copy
// of the previous level (slot >= 0); a row that reads a previous-level row
// past the LDS capacity is computed from memory instead (deferred)
template<class T, int MAXE>
struct LvlRow
{
    int32_t r, cnt, base, nLo, nEnt;
    bool deferred;        // reads a previous-level row past the LDS capacity
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
            const T value = sl >= 0 ? lds[sl] : o.v[p];
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

// stage 0 (three levels ahead): the row's shape
struct LvlS0
{
    int32_t r, info, base;
};

// stage 1 (two levels ahead): scalars, columns and coefficients in the
// update's order
template<class T, int MAXE>
struct LvlS1
{
    int32_t r, info, base;
    T a, dg;
    int32_t col[MAXE];
    T val[MAXE];
};

template<class T, int KIND>
__device__ __forceinline__ void lvl_s0(const RowsView& R, int32_t r, bool ok, LvlS0& o)
{
    o.r = ok ? r : -1;
    const int32_t rr = ok ? r : 0;
    o.info = ok ? R.rowInfo[rr] : 0;
    o.base = lvl_base(R.sliceStart, rr);
}

template<class T, int KIND, int MAXE>
__device__ __forceinline__ void lvl_s1(const RowsView& R, const SweepData<T>& d, const LvlS0& i,
                                       LvlS1<T, MAXE>& o)
{
    o.r = i.r;
    o.info = i.info;
    o.base = i.base;
    const int32_t rr = i.r >= 0 ? i.r : 0;
    if constexpr (KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD)
    {
        o.a = d.x[rr];
        o.dg = d.diag[rr];
    }
    else if constexpr (KIND == SWEEP_PRECOND_FWD)
    {
        o.a = d.x[rr];
        o.dg = d.rD[rr];
    }
    else
    {
        o.a = d.w[rr];         // the forward pass's value of the row
        o.dg = d.rD[rr];
    }
    const int32_t nLo = i.info & 0xffff, nEnt = i.info >> 16;
    const int32_t cnt = lvl_count<KIND>(nLo, nEnt);
    // every load is issued (an unused entry reads entry 0: the kernel runs
    // on meshes with faces only),
    // so no branch separates a load from its use a level later
#pragma unroll
    for (int p = 0; p < MAXE; p++)
    {
        const bool use = p < cnt;
        const int32_t e = use ? i.base + (lvl_entry<KIND>(p, nEnt) << 6) : 0;
        const int32_t c = R.col[e];
        const T v = d.val[e];
        o.col[p] = use ? c : -1;
        o.val[p] = use ? v : T(0);
    }
}

// stage 2 (one level ahead, at the start of the level before): the values,
// or the LDS slot of a value of that level, still being computed
// ([cLo, cHi)); values of the level finished last ([pLo, pHi)) from its LDS
// copy, older ones from the vector
template<class T, int KIND, int MAXE, int CAP>
__device__ __forceinline__ void lvl_s2(const SweepData<T>& d, const LvlS1<T, MAXE>& i,
                                       int32_t cLo, int32_t cHi, int32_t pLo, int32_t pHi,
                                       const T* ldsP, LvlRow<T, MAXE>& o)
{
    o.r = i.r;
    o.nLo = i.info & 0xffff;
    o.nEnt = i.info >> 16;
    o.cnt = i.r >= 0 ? lvl_count<KIND>(o.nLo, o.nEnt) : 0;
    o.base = i.base;
    o.a = i.a;
    o.dg = i.dg;
    o.deferred = false;
#pragma unroll
    for (int p = 0; p < MAXE; p++)
    {
        const int32_t c = i.col[p];
        const bool use = p < o.cnt;
        const bool inCur = use && c >= cLo && c < cHi;
        const bool inPrev = use && c >= pLo && c < pHi && c - pLo < CAP;
        const bool inVec = use && !inCur && !inPrev;
        // the vector read is issued unconditionally (row 0 when unused)
        T fromLds = T(0);
        if (inPrev) fromLds = ldsP[c - pLo];
        const T fromVec = d.w[inVec ? c : 0];
        o.val[p] = i.val[p];
        o.slot[p] = inCur && c - cLo < CAP ? c - cLo : -1;
        o.deferred = o.deferred || (inCur && c - cLo >= CAP);
        o.v[p] = inPrev ? fromLds : fromVec;
    }
}

// rows [lo, hi) of the level taken s-th in sweep order (empty outside)
struct LvlRange
{
    const int32_t* seg;
    int32_t nSeg;
    bool bwd;
    __device__ int32_t lo(int32_t s) const
    {
        return s < 0 || s >= nSeg ? 0 : seg[bwd ? nSeg - 1 - s : s];
    }
    __device__ int32_t hi(int32_t s) const
    {
        return s < 0 || s >= nSeg ? 0 : seg[(bwd ? nSeg - 1 - s : s) + 1];
    }
};

// one level: values of level s+1 (into `next`), columns of s+2 (s1),
// shape of s+3 (s0), then the rows of level s (from `cur`) and a barrier
template<class T, int KIND, int MAXE, int CAP>
__device__ __forceinline__ void lvl_level(const RowsView& R, const SweepData<T>& d,
                                          const LvlRange& G, int32_t s, T (*lv)[CAP],
                                          LvlRow<T, MAXE>& cur, LvlRow<T, MAXE>& next,
                                          LvlS1<T, MAXE>& s1, LvlS0& s0)
{
    const int32_t tid = (int32_t)threadIdx.x, nt = (int32_t)blockDim.x;
    const int32_t r0 = G.lo(s), r1 = G.hi(s);
    const int32_t pLo = G.lo(s - 1), pHi = G.hi(s - 1);
    const T* ldsPrev = lv[(s + 1) & 1];
    T* ldsMine = lv[s & 1];
    lvl_s2<T, KIND, MAXE, CAP>(d, s1, r0, r1, pLo, pHi, ldsPrev, next);
    lvl_s1<T, KIND, MAXE>(R, d, s0, s1);
    {
        const int32_t r = G.lo(s + 3) + tid;
        lvl_s0<T, KIND>(R, r, r < G.hi(s + 3), s0);
    }
    if (cur.r >= 0)
    {
        const T v = cur.deferred
                  ? lvl_direct<T, KIND, CAP>(R, d, cur.r, pLo, pHi, ldsPrev)
                  : lvl_compute<T, KIND, MAXE, CAP>(R, d, cur, pLo, pHi, ldsPrev);
        d.w[cur.r] = v;
        if (cur.r - r0 < CAP) ldsMine[cur.r - r0] = v;
    }
    for (int32_t r = r0 + tid + nt; r < r1; r += nt)
    {
        const T v = lvl_direct<T, KIND, CAP>(R, d, r, pLo, pHi, ldsPrev);
        d.w[r] = v;
        if (r - r0 < CAP) ldsMine[r - r0] = v;
    }
    __syncthreads();
}

template<class T, int KIND>
__global__ __launch_bounds__(LVL_THREADS)
void k_sweep_level_wg(RowsView R, SchedView S, SweepData<T> d, DevState* st)
{
    constexpr int CAP = LvlCfg<T>::CAP, MAXE = LvlCfg<T>::MAXE;
    __shared__ T lv[2][CAP];
    if (st && st->done) return;
    LvlRange G;
    G.nSeg = S.partLevelOff[1] - S.partLevelOff[0] - 1;
    G.seg = S.segStart + S.partLevelOff[0];
    G.bwd = LvlKind<KIND>::bwd;
    const int32_t tid = (int32_t)threadIdx.x;
    // pipeline: A (level s+1, values), s1 (level s+2, columns and
    // coefficients), s0 (level s+3, shape); the rows of level s are
    // computed from B
    LvlRow<T, MAXE> A, B;
    LvlS1<T, MAXE> s1;
    LvlS0 s0;
    lvl_s0<T, KIND>(R, G.lo(0) + tid, G.lo(0) + tid < G.hi(0), s0);
    lvl_s1<T, KIND, MAXE>(R, d, s0, s1);
    lvl_s2<T, KIND, MAXE, CAP>(d, s1, 0, 0, 0, 0, lv[0], B);
    lvl_s0<T, KIND>(R, G.lo(1) + tid, G.lo(1) + tid < G.hi(1), s0);
    lvl_s1<T, KIND, MAXE>(R, d, s0, s1);
    lvl_s0<T, KIND>(R, G.lo(2) + tid, G.lo(2) + tid < G.hi(2), s0);
    int32_t s = 0;
    for (; s + 1 < G.nSeg; s += 2)
    {
        lvl_level<T, KIND, MAXE, CAP>(R, d, G, s, lv, B, A, s1, s0);
        lvl_level<T, KIND, MAXE, CAP>(R, d, G, s + 1, lv, A, B, s1, s0);
    }
    if (s < G.nSeg) lvl_level<T, KIND, MAXE, CAP>(R, d, G, s, lv, B, A, s1, s0);
}

template<class T, int KIND>
static void launch_kind(const RowsView& R, const SchedView& S, const SweepData<T>& d,
                        DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep_level_wg<T, KIND>), dim3(1), dim3(LVL_THREADS), 0, s, R, S, d,
                       st);
}

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
