// Pairwise GAMG agglomeration on the device.
//
// pairGAMGAgglomerate.C:123-301 walks the cells in order (forward or
// backward); a free cell pairs with its free neighbour across the heaviest
// face (the first such face of its cellFaces list on ties), a cell with no
// free neighbour joins the cluster across its heaviest face, and the coarse
// cells are numbered in the order the pairs are made.  The walk is the
// greedy matching of the faces ranked by (walk position of the face's
// earlier cell, weight descending, position in that cell's face list):
// at its turn a cell's earlier neighbours are all taken, so its candidate
// faces are exactly the untaken faces it is the earlier cell of.  Greedy
// matching by a fixed ranking is computed here in rounds: every undecided
// cell proposes its best face to an undecided neighbour, a face proposed by
// both of its cells is in the matching (no better face touches it), both
// cells are decided, and a cell with no undecided neighbour is decided
// unmatched.  Only the undecided neighbours of cells decided in a round can
// change their proposal, so each round works on that list.  The result is
// the sequential walk's, whatever order the threads run in; the number of
// rounds is the length of the longest chain of dependent decisions (about
// 2.3 rounds per cell layer of a lexicographic hex box).
//
// Coarse numbering, joins and the coarse face addressing
// (GAMGAgglomerateLduAddressing.C:32-239: coarse faces by coarse owner, and
// within an owner in order of first appearance among the fine faces) are
// prefix sums and per-coarse-cell walks over sorted face lists.
#include "fh_agglom.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>

namespace fh
{

#define AG_BLOCK 256
#define AG_SCAN 1024           // elements per scan block (256 threads x 4)

static unsigned int ag_grid(int64_t n)
{
    return (unsigned int)std::max<int64_t>(1, (n + AG_BLOCK - 1)/AG_BLOCK);
}

// a fixed grid for the work lists, whose length lives on the device: one
// workgroup per CU covers a round's list (tens of thousands of cells at
// 200^3) in one pass; 2048 workgroups took 36 us per accept launch
static const unsigned int kListGrid = 256;

// ---------------------------------------------------------------------------
// exclusive prefix sum
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(AG_BLOCK)
void k_scan_block(const int32_t* __restrict__ in, int32_t* __restrict__ out, int32_t n,
                  int32_t* __restrict__ blockSums)
{
    __shared__ int32_t s[AG_BLOCK];
    const int32_t t = (int32_t)threadIdx.x;
    const int64_t base = (int64_t)blockIdx.x*AG_SCAN + 4*t;
    int32_t v[4];
#pragma unroll
    for (int i = 0; i < 4; i++) v[i] = base + i < n ? in[base + i] : 0;
    const int32_t p1 = v[0], p2 = p1 + v[1], p3 = p2 + v[2], p4 = p3 + v[3];
    s[t] = p4;
    __syncthreads();
    for (int off = 1; off < AG_BLOCK; off <<= 1)
    {
        const int32_t x = t >= off ? s[t - off] : 0;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    const int32_t e = t ? s[t - 1] : 0;
    if (base < n) out[base] = e;
    if (base + 1 < n) out[base + 1] = e + p1;
    if (base + 2 < n) out[base + 2] = e + p2;
    if (base + 3 < n) out[base + 3] = e + p3;
    if (t == AG_BLOCK - 1) blockSums[blockIdx.x] = s[t];
}

__global__ __launch_bounds__(AG_BLOCK)
void k_scan_add(int32_t* __restrict__ out, int32_t n, const int32_t* __restrict__ blockOff)
{
    const int64_t base = (int64_t)blockIdx.x*AG_SCAN + 4*threadIdx.x;
    const int32_t a = blockOff[blockIdx.x];
#pragma unroll
    for (int i = 0; i < 4; i++)
        if (base + i < n) out[base + i] += a;
}

// out[n] = out[n-1] + in[n-1] (the total), or 0 for n = 0
__global__ void k_scan_total(const int32_t* __restrict__ in, int32_t* __restrict__ out, int32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = n ? out[n - 1] + in[n - 1] : 0;
}

void device_exclusive_scan(const int32_t* in, int32_t* out, int32_t n, hipStream_t s)
{
    if (n > 0)
    {
        const int32_t nb = (int32_t)(((int64_t)n + AG_SCAN - 1)/AG_SCAN);
        DBuf<int32_t> sums, offs;
        sums.alloc(nb);
        hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(AG_BLOCK), 0, s, in, out, n, sums.p);
        if (nb > 1)
        {
            offs.alloc((size_t)nb + 1);
            device_exclusive_scan(sums.p, offs.p, nb, s);
            hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(AG_BLOCK), 0, s, out, n, offs.p);
        }
        // the temporaries are freed below: wait for their last reader
        hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, in, out, n);
        FH_HIP(hipStreamSynchronize(s));
        return;
    }
    hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, s, in, out, n);
}

static int32_t read_int(const int32_t* p, hipStream_t s)
{
    int32_t v = 0;
    FH_HIP(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    FH_HIP(hipStreamSynchronize(s));
    return v;
}

template<class T>
__global__ __launch_bounds__(AG_BLOCK)
void k_fill(T* __restrict__ p, int64_t n, T v)
{
    const int64_t i = (int64_t)blockIdx.x*AG_BLOCK + threadIdx.x;
    if (i < n) p[i] = v;
}

template<class T>
static void fill(T* p, int64_t n, T v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill<T>, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, p, n, v);
}

// ---------------------------------------------------------------------------
// cell -> face lists in the reference's cellFaces order
// (pairGAMGAgglomerate.C:136-182): the faces a cell is the upper cell of,
// then those it is the lower cell of, each in face order
// ---------------------------------------------------------------------------

struct CellFaces
{
    DBuf<int32_t> off;      // [n+1]
    DBuf<int32_t> nUp;      // [n] faces with the cell as upper cell (the first group)
    DBuf<int32_t> face;     // [2 nF] face of each entry
    DBuf<int32_t> nb;       // [2 nF] the other cell of the entry's face
    DBuf<int32_t> rankLo;   // [nF] position of the face in its lower cell's list
    DBuf<int32_t> rankUp;   // [nF] ... in its upper cell's list
};

__global__ __launch_bounds__(AG_BLOCK)
void k_cf_count(int32_t nF, const int32_t* __restrict__ lower, const int32_t* __restrict__ upper,
                int32_t* __restrict__ deg, int32_t* __restrict__ nUp)
{
    const int32_t f = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (f >= nF) return;
    atomicAdd(&deg[lower[f]], 1);
    atomicAdd(&deg[upper[f]], 1);
    atomicAdd(&nUp[upper[f]], 1);
}

// entries in any order within the two groups (sorted afterwards)
__global__ __launch_bounds__(AG_BLOCK)
void k_cf_fill(int32_t nF, const int32_t* __restrict__ lower, const int32_t* __restrict__ upper,
               const int32_t* __restrict__ off, const int32_t* __restrict__ nUp,
               int32_t* __restrict__ cntLo, int32_t* __restrict__ cntUp,
               int32_t* __restrict__ face)
{
    const int32_t f = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (f >= nF) return;
    const int32_t u = upper[f], l = lower[f];
    face[off[u] + atomicAdd(&cntUp[u], 1)] = f;
    face[off[l] + nUp[l] + atomicAdd(&cntLo[l], 1)] = f;
}

__device__ __forceinline__ void ag_isort(int32_t* a, int32_t n)
{
    for (int32_t i = 1; i < n; i++)
    {
        const int32_t v = a[i];
        int32_t j = i - 1;
        while (j >= 0 && a[j] > v)
        {
            a[j + 1] = a[j];
            j--;
        }
        a[j + 1] = v;
    }
}

__global__ __launch_bounds__(AG_BLOCK)
void k_cf_sort(int32_t n, const int32_t* __restrict__ lower, const int32_t* __restrict__ upper,
               const int32_t* __restrict__ off, const int32_t* __restrict__ nUp,
               int32_t* __restrict__ face, int32_t* __restrict__ nb,
               int32_t* __restrict__ rankLo, int32_t* __restrict__ rankUp)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= n) return;
    const int32_t o = off[c], e = off[c + 1], u = nUp[c];
    ag_isort(face + o, u);
    ag_isort(face + o + u, e - o - u);
    for (int32_t k = o; k < e; k++)
    {
        const int32_t f = face[k];
        if (k < o + u)
        {
            nb[k] = lower[f];
            rankUp[f] = k - o;
        }
        else
        {
            nb[k] = upper[f];
            rankLo[f] = k - o;
        }
    }
}

static void build_cell_faces(int32_t n, int32_t nF, const int32_t* lower, const int32_t* upper,
                             CellFaces& C, hipStream_t s)
{
    DBuf<int32_t> deg, cntLo, cntUp;
    deg.alloc((size_t)n + 1);
    C.nUp.alloc(n);
    cntLo.alloc(n);
    cntUp.alloc(n);
    fill(deg.p, (int64_t)n + 1, 0, s);
    fill(C.nUp.p, n, 0, s);
    fill(cntLo.p, n, 0, s);
    fill(cntUp.p, n, 0, s);
    C.off.alloc((size_t)n + 1);
    C.face.alloc(std::max<size_t>(1, 2*(size_t)nF));
    C.nb.alloc(std::max<size_t>(1, 2*(size_t)nF));
    C.rankLo.alloc(std::max<size_t>(1, nF));
    C.rankUp.alloc(std::max<size_t>(1, nF));
    if (nF > 0)
        hipLaunchKernelGGL(k_cf_count, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, lower, upper,
                           deg.p, C.nUp.p);
    device_exclusive_scan(deg.p, C.off.p, n, s);
    if (nF > 0)
    {
        hipLaunchKernelGGL(k_cf_fill, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, lower, upper,
                           C.off.p, C.nUp.p, cntLo.p, cntUp.p, C.face.p);
        hipLaunchKernelGGL(k_cf_sort, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n, lower, upper,
                           C.off.p, C.nUp.p, C.face.p, C.nb.p, C.rankLo.p, C.rankUp.p);
    }
    FH_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// the matching rounds
// ---------------------------------------------------------------------------

#define AG_UNDECIDED (-1)
#define AG_UNMATCHED (-2)

struct MatchView
{
    int32_t n;
    int32_t fwd;
    const int32_t *off, *face, *nb;
    const int32_t *lower, *upper, *rankLo, *rankUp;
    const double* w;
    int32_t* state;     // AG_UNDECIDED, AG_UNMATCHED or the matched face
    int32_t* prop;      // the cell's best face to an undecided neighbour, or -1
};

// ranking of a face: walk position of its earlier cell (the lower cell
// forward, the upper cell backward), then weight descending, then position
// in the earlier cell's face list
__device__ __forceinline__ bool ag_better(const MatchView& V, int32_t f, int32_t g)
{
    const int32_t pf = V.fwd ? V.lower[f] : V.n - 1 - V.upper[f];
    const int32_t pg = V.fwd ? V.lower[g] : V.n - 1 - V.upper[g];
    if (pf != pg) return pf < pg;
    const double wf = V.w[f], wg = V.w[g];
    if (wf != wg) return wf > wg;
    const int32_t rf = V.fwd ? V.rankLo[f] : V.rankUp[f];
    const int32_t rg = V.fwd ? V.rankLo[g] : V.rankUp[g];
    return rf < rg;
}

// list == nullptr: every cell
__global__ __launch_bounds__(AG_BLOCK)
void k_match_propose(MatchView V, const int32_t* __restrict__ list,
                     const int32_t* __restrict__ count, int32_t* __restrict__ zeroCount)
{
    const int32_t tid = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (zeroCount && tid == 0) *zeroCount = 0;
    const int32_t total = list ? *count : V.n;
    const int32_t stride = (int32_t)(gridDim.x*AG_BLOCK);
    for (int32_t i = tid; i < total; i += stride)
    {
        const int32_t c = list ? list[i] : i;
        if (V.state[c] != AG_UNDECIDED) continue;
        int32_t best = -1;
        for (int32_t k = V.off[c]; k < V.off[c + 1]; k++)
        {
            if (V.state[V.nb[k]] != AG_UNDECIDED) continue;
            const int32_t f = V.face[k];
            if (best < 0 || ag_better(V, f, best)) best = f;
        }
        V.prop[c] = best;
    }
}

__device__ __forceinline__ void ag_push(const MatchView& V, int32_t x, int32_t skip,
                                        int32_t* stamp, int32_t round, int32_t* next,
                                        int32_t* nextCount)
{
    for (int32_t k = V.off[x]; k < V.off[x + 1]; k++)
    {
        const int32_t y = V.nb[k];
        if (y == skip || V.state[y] != AG_UNDECIDED) continue;
        if (atomicExch(&stamp[y], round) != round) next[atomicAdd(nextCount, 1)] = y;
    }
}

// decided counts every cell once (the state moves from undecided once)
__global__ __launch_bounds__(AG_BLOCK)
void k_match_accept(MatchView V, const int32_t* __restrict__ list,
                    const int32_t* __restrict__ count, int32_t* __restrict__ next,
                    int32_t* __restrict__ nextCount, int32_t* __restrict__ stamp, int32_t round,
                    int32_t* __restrict__ decided)
{
    const int32_t tid = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    const int32_t total = list ? *count : V.n;
    const int32_t stride = (int32_t)(gridDim.x*AG_BLOCK);
    for (int32_t i = tid; i < total; i += stride)
    {
        const int32_t c = list ? list[i] : i;
        if (V.state[c] != AG_UNDECIDED) continue;
        const int32_t f = V.prop[c];
        if (f < 0)
        {
            // no undecided neighbour: the cell stays unpaired
            if (atomicCAS(&V.state[c], AG_UNDECIDED, AG_UNMATCHED) == AG_UNDECIDED)
                atomicAdd(decided, 1);
            continue;
        }
        const int32_t o = V.lower[f] == c ? V.upper[f] : V.lower[f];
        if (V.prop[o] != f) continue;
        // both cells' best face: in the matching (set by both cells'
        // threads when both are listed, with the same value)
        if (atomicCAS(&V.state[c], AG_UNDECIDED, f) == AG_UNDECIDED) atomicAdd(decided, 1);
        if (atomicCAS(&V.state[o], AG_UNDECIDED, f) == AG_UNDECIDED) atomicAdd(decided, 1);
        ag_push(V, c, o, stamp, round, next, nextCount);
        ag_push(V, o, c, stamp, round, next, nextCount);
    }
}

// ---------------------------------------------------------------------------
// coarse numbering: pairs in walk order of their earlier cell, unpaired
// cells join the cluster across their heaviest face, cells without faces
// are singletons after the pairs (pairGAMGAgglomerate.C:214-300)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(AG_BLOCK)
void k_number_flags(int32_t n, int32_t fwd, const int32_t* __restrict__ state,
                    const int32_t* __restrict__ lower, const int32_t* __restrict__ upper,
                    const int32_t* __restrict__ off, int32_t* __restrict__ creatorAt,
                    int32_t* __restrict__ singleAt)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= n) return;
    const int32_t p = fwd ? c : n - 1 - c;
    const int32_t f = state[c];
    const int32_t earlier = f >= 0 ? (fwd ? lower[f] : upper[f]) : -1;
    creatorAt[p] = earlier == c ? 1 : 0;
    singleAt[p] = off[c + 1] == off[c] ? 1 : 0;
}

__global__ __launch_bounds__(AG_BLOCK)
void k_number_assign(MatchView V, const int32_t* __restrict__ pairAt,
                     const int32_t* __restrict__ singleAt, int32_t* __restrict__ cmap,
                     int32_t* __restrict__ joinTo)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= V.n) return;
    const int32_t n = V.n;
    const int32_t f = V.state[c];
    joinTo[c] = -1;
    if (f >= 0)
    {
        const int32_t e = V.fwd ? V.lower[f] : V.upper[f];
        cmap[c] = pairAt[V.fwd ? e : n - 1 - e];
        return;
    }
    const int32_t o = V.off[c], e = V.off[c + 1];
    if (o == e)
    {
        cmap[c] = pairAt[n] + singleAt[V.fwd ? c : n - 1 - c];
        return;
    }
    // heaviest face, the first of the list on ties (strictly greater wins)
    int32_t best = V.face[o];
    double bw = V.w[best];
    for (int32_t k = o + 1; k < e; k++)
    {
        const int32_t g = V.face[k];
        if (V.w[g] > bw)
        {
            best = g;
            bw = V.w[g];
        }
    }
    cmap[c] = -1;
    joinTo[c] = V.lower[best] == c ? V.upper[best] : V.lower[best];
}

// an unpaired cell takes the coarse cell its join target had at the cell's
// turn: a paired cell's pair, or (an earlier unpaired cell) what that cell
// took in turn; the chain walks to earlier cells only
__global__ __launch_bounds__(AG_BLOCK)
void k_number_joins(int32_t n, const int32_t* __restrict__ joinTo, int32_t* __restrict__ cmap,
                    int32_t* __restrict__ err)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= n || joinTo[c] < 0) return;
    int32_t t = joinTo[c];
    int32_t steps = 0;
    while (joinTo[t] >= 0)
    {
        t = joinTo[t];
        if (++steps > n)
        {
            atomicExch(err, 1);
            return;
        }
    }
    cmap[c] = cmap[t];
}

__global__ __launch_bounds__(AG_BLOCK)
void k_number_reverse(int32_t n, int32_t last, int32_t* __restrict__ cmap)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c < n) cmap[c] = last - cmap[c];
}

// ---------------------------------------------------------------------------
// coarse face addressing (GAMGAgglomerateLduAddressing.C:32-239) and the
// restricted face weights (restrictFaceField, GAMGAgglomerationTemplates.C:
// 137-167: summed in fine face order)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(AG_BLOCK)
void k_members_count(int32_t n, const int32_t* __restrict__ cmap,
                     const int32_t* __restrict__ off, int32_t* __restrict__ nMem,
                     int32_t* __restrict__ segLen)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= n) return;
    atomicAdd(&nMem[cmap[c]], 1);
    atomicAdd(&segLen[cmap[c]], off[c + 1] - off[c]);
}

__global__ __launch_bounds__(AG_BLOCK)
void k_members_fill(int32_t n, const int32_t* __restrict__ cmap,
                    const int32_t* __restrict__ memOff, int32_t* __restrict__ cnt,
                    int32_t* __restrict__ members)
{
    const int32_t c = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (c >= n) return;
    const int32_t C = cmap[c];
    members[memOff[C] + atomicAdd(&cnt[C], 1)] = c;
}

// per coarse cell C: the fine faces it owns (the other cell's coarse index
// above C) with that index, sorted by fine face; the number of distinct
// coarse neighbours
__global__ __launch_bounds__(AG_BLOCK)
void k_owned_faces(int32_t nC, const int32_t* __restrict__ cmap,
                   const int32_t* __restrict__ off, const int32_t* __restrict__ face,
                   const int32_t* __restrict__ nb, const int32_t* __restrict__ memOff,
                   const int32_t* __restrict__ members, const int32_t* __restrict__ segOff,
                   int32_t* __restrict__ segF, int32_t* __restrict__ segN,
                   int32_t* __restrict__ segCount, int32_t* __restrict__ nDistinct)
{
    const int32_t C = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (C >= nC) return;
    const int32_t s0 = segOff[C];
    int32_t m = 0;
    for (int32_t q = memOff[C]; q < memOff[C + 1]; q++)
    {
        const int32_t c = members[q];
        for (int32_t k = off[c]; k < off[c + 1]; k++)
        {
            const int32_t N = cmap[nb[k]];
            if (N <= C) continue;
            // insert by fine face (the member lists are each sorted)
            const int32_t f = face[k];
            int32_t j = m - 1;
            while (j >= 0 && segF[s0 + j] > f)
            {
                segF[s0 + j + 1] = segF[s0 + j];
                segN[s0 + j + 1] = segN[s0 + j];
                j--;
            }
            segF[s0 + j + 1] = f;
            segN[s0 + j + 1] = N;
            m++;
        }
    }
    segCount[C] = m;
    int32_t d = 0;
    for (int32_t i = 0; i < m; i++)
    {
        const int32_t N = segN[s0 + i];
        bool first = true;
        for (int32_t j = 0; j < i; j++)
            if (segN[s0 + j] == N) { first = false; break; }
        d += first ? 1 : 0;
    }
    nDistinct[C] = d;
}

__global__ __launch_bounds__(AG_BLOCK)
void k_coarse_faces(int32_t nC, const int32_t* __restrict__ segOff,
                    const int32_t* __restrict__ segF, const int32_t* __restrict__ segN,
                    int32_t* __restrict__ segId, const int32_t* __restrict__ segCount,
                    const int32_t* __restrict__ cfOff, const int32_t* __restrict__ upper,
                    const int32_t* __restrict__ cmap, const double* __restrict__ w,
                    int32_t* __restrict__ cLower, int32_t* __restrict__ cUpper,
                    double* __restrict__ cw, int32_t* __restrict__ faceRestrict,
                    uint8_t* __restrict__ faceFlip)
{
    const int32_t C = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (C >= nC) return;
    const int32_t s0 = segOff[C], m = segCount[C];
    int32_t d = cfOff[C];
    for (int32_t i = 0; i < m; i++)
    {
        const int32_t N = segN[s0 + i];
        int32_t id = -1;
        for (int32_t j = 0; j < i; j++)
            if (segN[s0 + j] == N) { id = segId[s0 + j]; break; }
        if (id < 0)
        {
            id = d++;
            cLower[id] = C;
            cUpper[id] = N;
            cw[id] = 0.0;
        }
        segId[s0 + i] = id;
        const int32_t f = segF[s0 + i];
        faceRestrict[f] = id;
        // flipped when the fine upper cell lands on the coarse owner
        faceFlip[f] = cmap[upper[f]] == C ? 1 : 0;
        cw[id] += w[f];
    }
}

// faces inside a coarse cell restrict to -1 - that cell
__global__ __launch_bounds__(AG_BLOCK)
void k_inner_faces(int32_t nF, const int32_t* __restrict__ lower,
                   const int32_t* __restrict__ upper, const int32_t* __restrict__ cmap,
                   int32_t* __restrict__ faceRestrict, uint8_t* __restrict__ faceFlip)
{
    const int32_t f = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (f >= nF) return;
    const int32_t a = cmap[lower[f]];
    if (a == cmap[upper[f]])
    {
        faceRestrict[f] = -a - 1;
        faceFlip[f] = 0;
    }
}

// combineLevels (GAMGAgglomerateLduAddressing.C:514-560): the previous
// level's maps composed with the new level's
__global__ __launch_bounds__(AG_BLOCK)
void k_combine_faces(int32_t nF, int32_t nMidFaces, int32_t* __restrict__ prevFace,
                     uint8_t* __restrict__ prevFlip, const int32_t* __restrict__ curFace,
                     const uint8_t* __restrict__ curFlip, const int32_t* __restrict__ curCell)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i >= nF) return;
    const int32_t r = prevFace[i];
    if (r >= 0)
    {
        prevFace[i] = curFace[r];
        prevFlip[i] = curFlip[r];
    }
    else
    {
        // the reference reads the flip map at the cell index here; a face
        // inside a coarse cell goes to the diagonal, where the flip is not used
        const int32_t c = -r - 1;
        prevFace[i] = -curCell[c] - 1;
        prevFlip[i] = c < nMidFaces ? curFlip[c] : 0;
    }
}

__global__ __launch_bounds__(AG_BLOCK)
void k_combine_cells(int32_t n, int32_t* __restrict__ prevCell, const int32_t* __restrict__ curCell)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i < n) prevCell[i] = curCell[prevCell[i]];
}

// algebraicPairGAMGAgglomeration.C:57-64: mag(upper), or for an asymmetric
// matrix max(mag(upper), mag(lower))
__global__ __launch_bounds__(AG_BLOCK)
void k_agg_weights(int32_t nF, const double* __restrict__ upperF,
                   const double* __restrict__ lowerF, double* __restrict__ w)
{
    const int32_t f = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (f >= nF) return;
    const double u = fabs(upperF[f]);
    w[f] = lowerF ? fmax(u, fabs(lowerF[f])) : u;
}

void launch_agg_weights(int32_t nF, const double* upperF, const double* lowerF, double* w,
                        hipStream_t s)
{
    if (nF > 0)
        hipLaunchKernelGGL(k_agg_weights, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, upperF,
                           lowerF, w);
}

// ---------------------------------------------------------------------------
// grouped lists (the per-level transfer lists of the V-cycle)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(AG_BLOCK)
void k_bucket_count(int32_t nItems, const int32_t* __restrict__ key, int32_t* __restrict__ cnt)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i < nItems && key[i] >= 0) atomicAdd(&cnt[key[i]], 1);
}

__global__ __launch_bounds__(AG_BLOCK)
void k_bucket_rows(int32_t nRows, const int32_t* __restrict__ rowKey,
                   const int32_t* __restrict__ cnt, int32_t* __restrict__ cntRow,
                   int32_t* __restrict__ keyRow)
{
    const int32_t r = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (r >= nRows) return;
    const int32_t k = rowKey ? rowKey[r] : r;
    cntRow[r] = cnt[k];
    keyRow[k] = r;
}

__global__ __launch_bounds__(AG_BLOCK)
void k_bucket_fill(int32_t nItems, const int32_t* __restrict__ key,
                   const int32_t* __restrict__ keyRow, const int32_t* __restrict__ start,
                   int32_t* __restrict__ fillCnt, int32_t* __restrict__ list)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i >= nItems || key[i] < 0) return;
    const int32_t r = keyRow[key[i]];
    list[start[r] + atomicAdd(&fillCnt[r], 1)] = i;
}

__global__ __launch_bounds__(AG_BLOCK)
void k_bucket_sort(int32_t nRows, const int32_t* __restrict__ start, int32_t* __restrict__ list)
{
    const int32_t r = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (r < nRows) ag_isort(list + start[r], start[r + 1] - start[r]);
}

void device_bucket(int32_t nItems, const int32_t* key, int32_t nRows, const int32_t* rowKey,
                   DBuf<int32_t>& start, DBuf<int32_t>& list, hipStream_t s)
{
    DBuf<int32_t> cnt, cntRow, keyRow, fillCnt;
    cnt.alloc(std::max(nRows, 1));
    cntRow.alloc(std::max(nRows, 1));
    keyRow.alloc(std::max(nRows, 1));
    fillCnt.alloc(std::max(nRows, 1));
    fill(cnt.p, nRows, 0, s);
    fill(fillCnt.p, nRows, 0, s);
    if (nItems > 0)
        hipLaunchKernelGGL(k_bucket_count, dim3(ag_grid(nItems)), dim3(AG_BLOCK), 0, s, nItems,
                           key, cnt.p);
    if (nRows > 0)
        hipLaunchKernelGGL(k_bucket_rows, dim3(ag_grid(nRows)), dim3(AG_BLOCK), 0, s, nRows,
                           rowKey, cnt.p, cntRow.p, keyRow.p);
    start.alloc((size_t)nRows + 1);
    device_exclusive_scan(cntRow.p, start.p, nRows, s);
    const int32_t total = read_int(start.p + nRows, s);
    list.alloc(std::max(total, 1));
    if (nItems > 0)
        hipLaunchKernelGGL(k_bucket_fill, dim3(ag_grid(nItems)), dim3(AG_BLOCK), 0, s, nItems,
                           key, keyRow.p, start.p, fillCnt.p, list.p);
    if (nRows > 0)
        hipLaunchKernelGGL(k_bucket_sort, dim3(ag_grid(nRows)), dim3(AG_BLOCK), 0, s, nRows,
                           start.p, list.p);
    FH_HIP(hipStreamSynchronize(s));
}

__global__ __launch_bounds__(AG_BLOCK)
void k_inverse_perm(int32_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ iperm)
{
    const int32_t r = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (r < n) iperm[perm[r]] = r;
}

// out[i] = map[in[i]]
__global__ __launch_bounds__(AG_BLOCK)
void k_gather_map(int32_t n, const int32_t* __restrict__ in, const int32_t* __restrict__ map,
                  int32_t* __restrict__ out)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i < n) out[i] = map[in[i]];
}

__global__ __launch_bounds__(AG_BLOCK)
void k_gather_u8(int32_t n, const int32_t* __restrict__ in, const uint8_t* __restrict__ src,
                 uint8_t* __restrict__ out)
{
    const int32_t i = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (i < n) out[i] = src[in[i]];
}

// keys of the fine faces: the coarse cell of a face inside one (inner) or
// the coarse face (outer), -1 otherwise
__global__ __launch_bounds__(AG_BLOCK)
void k_face_keys(int32_t nF, const int32_t* __restrict__ faceRestrict, int32_t inner,
                 int32_t* __restrict__ key)
{
    const int32_t f = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (f >= nF) return;
    const int32_t r = faceRestrict[f];
    key[f] = inner ? (r < 0 ? -r - 1 : -1) : (r >= 0 ? r : -1);
}

__global__ __launch_bounds__(AG_BLOCK)
void k_prolong_rows(int32_t nFine, const int32_t* __restrict__ finePerm,
                    const int32_t* __restrict__ restrictMap,
                    const int32_t* __restrict__ coarseIperm, int32_t* __restrict__ prolong)
{
    const int32_t r = (int32_t)(blockIdx.x*AG_BLOCK + threadIdx.x);
    if (r < nFine) prolong[r] = coarseIperm[restrictMap[finePerm[r]]];
}

void device_transfer_lists(const AggMaps& M, const int32_t* finePerm, const int32_t* coarsePerm,
                           TransferLists& T, hipStream_t s)
{
    const int32_t nF = M.nFine, nC = M.nCoarse, nFF = M.nFineFaces, nCF = M.nCoarseFaces;
    DBuf<int32_t> fineIperm, coarseIperm, key;
    fineIperm.alloc(std::max(nF, 1));
    coarseIperm.alloc(std::max(nC, 1));
    hipLaunchKernelGGL(k_inverse_perm, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, finePerm,
                       fineIperm.p);
    hipLaunchKernelGGL(k_inverse_perm, dim3(ag_grid(nC)), dim3(AG_BLOCK), 0, s, nC, coarsePerm,
                       coarseIperm.p);
    // fine cells of each coarse row, ascending fine natural order, as fine rows
    device_bucket(nF, M.restrictMap.p, nC, coarsePerm, T.rStart, T.rCells, s);
    if (nF > 0)
        hipLaunchKernelGGL(k_gather_map, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, T.rCells.p,
                           fineIperm.p, T.rCells.p);
    T.prolong.alloc(std::max(nF, 1));
    hipLaunchKernelGGL(k_prolong_rows, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, finePerm,
                       M.restrictMap.p, coarseIperm.p, T.prolong.p);
    // fine faces inside each coarse row's cell, ascending
    key.alloc(std::max(nFF, 1));
    if (nFF > 0)
        hipLaunchKernelGGL(k_face_keys, dim3(ag_grid(nFF)), dim3(AG_BLOCK), 0, s, nFF,
                           M.faceRestrict.p, 1, key.p);
    device_bucket(nFF, key.p, nC, coarsePerm, T.iStart, T.iFaces, s);
    // fine faces of each coarse face (coarse face order), ascending, with flips
    if (nFF > 0)
        hipLaunchKernelGGL(k_face_keys, dim3(ag_grid(nFF)), dim3(AG_BLOCK), 0, s, nFF,
                           M.faceRestrict.p, 0, key.p);
    device_bucket(nFF, key.p, nCF, nullptr, T.fStart, T.fFaces, s);
    const int32_t nOuter = read_int(T.fStart.p + nCF, s);
    T.fFlip.alloc(std::max(nOuter, 1));
    if (nOuter > 0)
        hipLaunchKernelGGL(k_gather_u8, dim3(ag_grid(nOuter)), dim3(AG_BLOCK), 0, s, nOuter,
                           T.fFaces.p, M.faceFlip.p, T.fFlip.p);
    FH_HIP(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// one pair level
// ---------------------------------------------------------------------------

// rounds between two reads of the list length
static const int32_t kRoundBatch = 32;

static std::unique_ptr<AggMaps> pair_level(int32_t n, int32_t nF, const int32_t* lower,
                                           const int32_t* upper, const double* w, bool fwd,
                                           hipStream_t s)
{
    static const bool verbose = std::getenv("FOAMHIP_GAMG_VERBOSE") != nullptr;
    CellFaces CF;
    build_cell_faces(n, nF, lower, upper, CF, s);

    DBuf<int32_t> state, prop, lists, counts, stamp;
    state.alloc(std::max(n, 1));
    prop.alloc(std::max(n, 1));
    lists.alloc(std::max<size_t>(2, 2*(size_t)n));
    counts.alloc(3);                 // the two list lengths, decided cells
    stamp.alloc(std::max(n, 1));
    fill(state.p, n, (int32_t)AG_UNDECIDED, s);
    fill(prop.p, n, -1, s);
    fill(stamp.p, n, -1, s);
    fill(counts.p, 3, 0, s);
    MatchView V{n, fwd ? 1 : 0, CF.off.p, CF.face.p, CF.nb.p, lower, upper, CF.rankLo.p,
                CF.rankUp.p, w, state.p, prop.p};
    const unsigned int gAll = std::min<unsigned int>(ag_grid(n), kListGrid);
    // round 1 over every cell; round r reads list (r-1)&1 and writes list r&1
    hipLaunchKernelGGL(k_match_propose, dim3(gAll), dim3(AG_BLOCK), 0, s, V,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, counts.p + 1);
    hipLaunchKernelGGL(k_match_accept, dim3(gAll), dim3(AG_BLOCK), 0, s, V,
                       (const int32_t*)nullptr, (const int32_t*)nullptr, lists.p + n,
                       counts.p + 1, stamp.p, 1, counts.p + 2);
    int32_t round = 1;
    // every round decides at least the cells of the best remaining face, so
    // a batch of rounds without a decision is a broken invariant, reported
    // instead of looping
    int32_t decidedBefore = -1;
    for (;;)
    {
        const int32_t decided = read_int(counts.p + 2, s);
        if (decided >= n) break;
        if (decided == decidedBefore || read_int(counts.p + (round & 1), s) == 0)
            fail(FH_ERR_INVALID, "internal: the agglomeration rounds stopped deciding cells");
        decidedBefore = decided;
        for (int32_t b = 0; b < kRoundBatch; b++)
        {
            round++;
            const int32_t cur = (round - 1) & 1, nxt = round & 1;
            hipLaunchKernelGGL(k_match_propose, dim3(kListGrid), dim3(AG_BLOCK), 0, s, V,
                               lists.p + (size_t)cur*n, counts.p + cur, counts.p + nxt);
            hipLaunchKernelGGL(k_match_accept, dim3(kListGrid), dim3(AG_BLOCK), 0, s, V,
                               lists.p + (size_t)cur*n, counts.p + cur, lists.p + (size_t)nxt*n,
                               counts.p + nxt, stamp.p, round, counts.p + 2);
        }
    }
    if (verbose)
        std::fprintf(stderr, "foamhip GAMG pair level: %d cells, %d faces, %d matching rounds\n",
                     n, nF, round);

    // numbering
    DBuf<int32_t> creatorAt, singleAt, pairAt, singleRank, joinTo, err;
    creatorAt.alloc(n);
    singleAt.alloc(n);
    pairAt.alloc((size_t)n + 1);
    singleRank.alloc((size_t)n + 1);
    joinTo.alloc(n);
    err.alloc(1);
    fill(err.p, 1, 0, s);
    auto M = std::make_unique<AggMaps>();
    M->nFine = n;
    M->nFineFaces = nF;
    M->restrictMap.alloc(n);
    hipLaunchKernelGGL(k_number_flags, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n, V.fwd, state.p,
                       lower, upper, CF.off.p, creatorAt.p, singleAt.p);
    device_exclusive_scan(creatorAt.p, pairAt.p, n, s);
    device_exclusive_scan(singleAt.p, singleRank.p, n, s);
    hipLaunchKernelGGL(k_number_assign, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, V, pairAt.p,
                       singleRank.p, M->restrictMap.p, joinTo.p);
    hipLaunchKernelGGL(k_number_joins, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n, joinTo.p,
                       M->restrictMap.p, err.p);
    const int32_t nC = read_int(pairAt.p + n, s) + read_int(singleRank.p + n, s);
    if (read_int(err.p, s)) fail(FH_ERR_INVALID, "internal: agglomeration join chain");
    if (!fwd)
        hipLaunchKernelGGL(k_number_reverse, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n, nC - 1,
                           M->restrictMap.p);
    M->nCoarse = nC;

    // coarse faces
    DBuf<int32_t> nMem, segLen, memOff, cnt, members, segOff, segF, segN, segId, segCount,
                  nDistinct, cfOff;
    nMem.alloc((size_t)nC + 1);
    segLen.alloc((size_t)nC + 1);
    cnt.alloc(nC);
    fill(nMem.p, (int64_t)nC + 1, 0, s);
    fill(segLen.p, (int64_t)nC + 1, 0, s);
    fill(cnt.p, nC, 0, s);
    hipLaunchKernelGGL(k_members_count, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n,
                       M->restrictMap.p, CF.off.p, nMem.p, segLen.p);
    memOff.alloc((size_t)nC + 1);
    segOff.alloc((size_t)nC + 1);
    device_exclusive_scan(nMem.p, memOff.p, nC, s);
    device_exclusive_scan(segLen.p, segOff.p, nC, s);
    members.alloc(n);
    hipLaunchKernelGGL(k_members_fill, dim3(ag_grid(n)), dim3(AG_BLOCK), 0, s, n,
                       M->restrictMap.p, memOff.p, cnt.p, members.p);
    const size_t segTotal = std::max<size_t>(1, 2*(size_t)nF);
    segF.alloc(segTotal);
    segN.alloc(segTotal);
    segId.alloc(segTotal);
    segCount.alloc(nC);
    nDistinct.alloc(nC);
    hipLaunchKernelGGL(k_owned_faces, dim3(ag_grid(nC)), dim3(AG_BLOCK), 0, s, nC,
                       M->restrictMap.p, CF.off.p, CF.face.p, CF.nb.p, memOff.p, members.p,
                       segOff.p, segF.p, segN.p, segCount.p, nDistinct.p);
    cfOff.alloc((size_t)nC + 1);
    device_exclusive_scan(nDistinct.p, cfOff.p, nC, s);
    const int32_t nCF = read_int(cfOff.p + nC, s);
    M->nCoarseFaces = nCF;
    M->lower.alloc(std::max(nCF, 1));
    M->upper.alloc(std::max(nCF, 1));
    M->weights.alloc(std::max(nCF, 1));
    M->faceRestrict.alloc(std::max(nF, 1));
    M->faceFlip.alloc(std::max(nF, 1));
    hipLaunchKernelGGL(k_coarse_faces, dim3(ag_grid(nC)), dim3(AG_BLOCK), 0, s, nC, segOff.p,
                       segF.p, segN.p, segId.p, segCount.p, cfOff.p, upper, M->restrictMap.p, w,
                       M->lower.p, M->upper.p, M->weights.p, M->faceRestrict.p, M->faceFlip.p);
    if (nF > 0)
        hipLaunchKernelGGL(k_inner_faces, dim3(ag_grid(nF)), dim3(AG_BLOCK), 0, s, nF, lower,
                           upper, M->restrictMap.p, M->faceRestrict.p, M->faceFlip.p);
    FH_HIP(hipStreamSynchronize(s));
    return M;
}

// a takes b's buffer (a's own is freed)
template<class T>
static void take(DBuf<T>& a, DBuf<T>& b)
{
    a.release();
    a.p = b.p;
    a.n = b.n;
    b.p = nullptr;
    b.n = 0;
}

std::vector<std::unique_ptr<AggMaps>> device_agglomerate(int32_t nCells, int32_t nFaces,
                                                         const int32_t* dLower,
                                                         const int32_t* dUpper,
                                                         const double* dWeights,
                                                         int32_t minCells, int32_t mergeLevels,
                                                         bool& forward, hipStream_t s)
{
    // pairGAMGAgglomerate.C:31-118; maxLevels_ 50 (GAMGAgglomeration.C:283)
    const int maxLevels = 50;
    if (mergeLevels < 1) mergeLevels = 1;
    std::vector<std::unique_ptr<AggMaps>> levels;
    const int32_t* lo = dLower;
    const int32_t* up = dUpper;
    const double* w = dWeights;
    int32_t n = nCells, nF = nFaces;
    int nPairLevels = 0;
    while ((int)levels.size() < maxLevels - 1)
    {
        std::unique_ptr<AggMaps> L = pair_level(n, nF, lo, up, w, forward, s);
        forward = !forward;
        // continueAgglomerating (GAMGAgglomeration.C:205-230)
        if (!(L->nCoarse >= minCells && L->nCoarse < n)) break;
        n = L->nCoarse;
        nF = L->nCoarseFaces;
        if (nPairLevels % mergeLevels)
        {
            AggMaps& P = *levels.back();
            if (P.nFineFaces > 0)
                hipLaunchKernelGGL(k_combine_faces, dim3(ag_grid(P.nFineFaces)), dim3(AG_BLOCK),
                                   0, s, P.nFineFaces, L->nFineFaces, P.faceRestrict.p,
                                   P.faceFlip.p, L->faceRestrict.p, L->faceFlip.p,
                                   L->restrictMap.p);
            hipLaunchKernelGGL(k_combine_cells, dim3(ag_grid(P.nFine)), dim3(AG_BLOCK), 0, s,
                               P.nFine, P.restrictMap.p, L->restrictMap.p);
            FH_HIP(hipStreamSynchronize(s));
            P.nCoarse = L->nCoarse;
            P.nCoarseFaces = L->nCoarseFaces;
            take(P.lower, L->lower);
            take(P.upper, L->upper);
            take(P.weights, L->weights);
            lo = P.lower.p;
            up = P.upper.p;
            w = P.weights.p;
        }
        else
        {
            lo = L->lower.p;
            up = L->upper.p;
            w = L->weights.p;
            levels.push_back(std::move(L));
        }
        nPairLevels++;
    }
    return levels;
}

} // namespace fh
