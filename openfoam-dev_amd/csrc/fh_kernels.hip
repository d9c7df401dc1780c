// HIP kernels of the foamhip lduMatrix solver library, written for gfx950
// (CDNA4, 64-wide wavefronts).  Every kernel that reproduces a reference
// loop keeps that loop's per-row accumulation order; the library is built
// with -ffp-contract=off so no multiply-add is contracted into an FMA.
//
// Rows are in the library numbering (see fh_mesh.cpp); the off-diagonal
// entries of row r are stored SELL-64: slice r/64, entry k of row r at
// sliceStart[r/64] + 64*k + r%64, so one lane per row reads coalesced.
#include "fh_kernels.h"

#include <cfloat>
#include <cstdint>

namespace fh
{

#define FH_BLOCK 256

__device__ __forceinline__ int32_t sell_base(const int32_t* sliceStart, int32_t r)
{
    return sliceStart[r >> 6] + (r & 63);
}

// --------------------------------------------------------------------------
// deterministic block reduction: xor-butterfly within each wave, then the
// wave sums in wave order.  Result valid in thread 0.
// --------------------------------------------------------------------------
template<int NT>
__device__ __forceinline__ double block_sum(double v, double* smem)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
    {
#pragma unroll
        for (int i = 0; i < NT/64; i++) s += smem[i];
    }
    return s;
}

int32_t reduction_blocks(int32_t n) { return (n + FH_BLOCK - 1)/FH_BLOCK; }
size_t reduction_partials(int32_t n) { return (size_t)reduction_blocks(n) + 1; }

// --------------------------------------------------------------------------
// coefficient gather: SELL values from the face-ordered upper/lower arrays
// (lower entry of row u from face f is lower[f], upper entry is upper[f],
// lduMatrixATmul.C:73-79)
// --------------------------------------------------------------------------
template<class T>
__global__ void k_gather_values(int64_t nnzPad, const int32_t* __restrict__ face,
                                const double* __restrict__ upperF,
                                const double* __restrict__ lowerF, T* __restrict__ val)
{
    const int64_t e = (int64_t)blockIdx.x*FH_BLOCK + threadIdx.x;
    if (e >= nnzPad) return;
    const int32_t fc = face[e];
    double v = 0.0;
    if (fc >= 0)
    {
        const int32_t f = fc >> 1;
        v = (fc & 1) ? upperF[f] : lowerF[f];
    }
    val[e] = (T)v;
}

template<class T>
void launch_gather_values(const RowsView& R, int64_t nnzPad, const double* upperF,
                          const double* lowerF, T* val, hipStream_t s)
{
    if (nnzPad == 0) return;
    const int64_t nb = (nnzPad + FH_BLOCK - 1)/FH_BLOCK;
    hipLaunchKernelGGL(k_gather_values<T>, dim3((unsigned)nb), dim3(FH_BLOCK), 0, s,
                       nnzPad, R.face, upperF, lowerF, val);
}

template<class Tin, class Tout>
__global__ void k_permute_in(int32_t n, const int32_t* __restrict__ perm,
                             const Tin* __restrict__ xNat, Tout* __restrict__ xNew)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r < n) xNew[r] = (Tout)xNat[perm[r]];
}

template<class Tin, class Tout>
__global__ void k_permute_out(int32_t n, const int32_t* __restrict__ perm,
                              const Tin* __restrict__ xNew, Tout* __restrict__ xNat)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r < n) xNat[perm[r]] = (Tout)xNew[r];
}

template<class Tin, class Tout>
void launch_permute_in(int32_t n, const int32_t* perm, const Tin* xNat, Tout* xNew,
                       hipStream_t s)
{
    hipLaunchKernelGGL((k_permute_in<Tin, Tout>), dim3(reduction_blocks(n)), dim3(FH_BLOCK),
                       0, s, n, perm, xNat, xNew);
}

template<class Tin, class Tout>
void launch_permute_out(int32_t n, const int32_t* perm, const Tin* xNew, Tout* xNat,
                        hipStream_t s)
{
    hipLaunchKernelGGL((k_permute_out<Tin, Tout>), dim3(reduction_blocks(n)), dim3(FH_BLOCK),
                       0, s, n, perm, xNew, xNat);
}

template<class Tin, class Tout>
__global__ void k_convert(int64_t n, const Tin* __restrict__ a, Tout* __restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) b[i] = (Tout)a[i];
}

template<class Tin, class Tout>
void launch_convert(int64_t n, const Tin* a, Tout* b, hipStream_t s)
{
    if (n == 0) return;
    hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3((unsigned)((n + FH_BLOCK - 1)/FH_BLOCK)),
                       dim3(FH_BLOCK), 0, s, n, a, b);
}

// --------------------------------------------------------------------------
// Amul / Tmul / residual / sumA
// --------------------------------------------------------------------------

// lduMatrixATmul.C:34-92: Apsi[c] = diag*psi, then += coefficient*psi(nbr)
// over the row's faces in face order
template<class T>
__global__ void k_amul(RowsView R, const T* __restrict__ diag, const T* __restrict__ val,
                       const T* __restrict__ x, T* __restrict__ y, const DevState* st)
{
    if (st && st->done) return;
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = sell_base(R.sliceStart, r);
    T acc = diag[r]*x[r];
    for (int32_t k = 0; k < nEnt; k++)
    {
        const int32_t e = base + (k << 6);
        acc += val[e]*x[R.col[e]];
    }
    // coupled interfaces (lduMatrixATmul.C:81-88, updateMatrixInterfaces):
    // result -= coeff*psi(nbr) with the stored value -coeff
    y[r] = add_interfaces(R, r, val, x, acc);
}

template<class T>
void launch_amul(const RowsView& R, const T* diag, const T* val, const T* x, T* y,
                 const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_amul<T>, dim3(reduction_blocks(R.n)), dim3(FH_BLOCK), 0, s,
                       R, diag, val, x, y, st);
}

// lduMatrixATmul.C:95-151: the same walk with upper/lower swapped, i.e. the
// SELL values of the transposed roles (built by the caller)
template<class T>
void launch_tmul(const RowsView& R, const T* diag, const T* valT, const T* x, T* y,
                 hipStream_t s)
{
    launch_amul<T>(R, diag, valT, x, y, nullptr, s);
}

// lduMatrixATmul.C:203-280: rA = source - diag*psi, then -= coeff*psi(nbr)
template<class T>
__global__ void k_residual(RowsView R, const T* __restrict__ diag, const T* __restrict__ val,
                           const T* __restrict__ x, const T* __restrict__ b,
                           T* __restrict__ res, const DevState* st)
{
    if (st && st->done) return;
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = sell_base(R.sliceStart, r);
    T acc = b[r] - diag[r]*x[r];
    for (int32_t k = 0; k < nEnt; k++)
    {
        const int32_t e = base + (k << 6);
        acc -= val[e]*x[R.col[e]];
    }
    // coupled interfaces with the negated coefficients (mBouCoeffs,
    // lduMatrixATmul.C:230-244, 266-273): rA -= mBou*psi(nbr), mBou the
    // stored value
    if (R.ifStart)
        for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++)
            acc -= val[R.ifBase + e]*x[R.ifNbr[e]];
    res[r] = acc;
}

template<class T>
void launch_residual(const RowsView& R, const T* diag, const T* val, const T* x,
                     const T* b, T* r, const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_residual<T>, dim3(reduction_blocks(R.n)), dim3(FH_BLOCK), 0, s,
                       R, diag, val, x, b, r, st);
}

// lduMatrixATmul.C:154-200: sumA = diag + sum of the row's coefficients
__global__ void k_sum_a(RowsView R, const double* __restrict__ diag,
                        const double* __restrict__ val, double* __restrict__ sumA)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    const int32_t nEnt = R.rowInfo[r] >> 16;
    const int32_t base = sell_base(R.sliceStart, r);
    double acc = diag[r];
    for (int32_t k = 0; k < nEnt; k++) acc += val[base + (k << 6)];
    // sumA[faceCells] -= interfaceBouCoeffs (lduMatrixATmul.C:185-199)
    if (R.ifStart)
        for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++) acc += val[R.ifBase + e];
    sumA[r] = acc;
}

// Gauss-Seidel with coupled interfaces: bPrime = source, then the
// interfaces as a Jacobi term of psi before the sweep with the negated
// coefficients (GaussSeidelSmoother.C:117-143): bPrime -= mBou*psi(nbr)
__global__ void k_interface_bprime(RowsView R, const double* __restrict__ val,
                                   const double* __restrict__ psi,
                                   const double* __restrict__ src, double* __restrict__ bPrime)
{
    const int32_t r = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (r >= R.n) return;
    double acc = src[r];
    for (int32_t e = R.ifStart[r]; e < R.ifStart[r + 1]; e++)
        acc -= val[R.ifBase + e]*psi[R.ifNbr[e]];
    bPrime[r] = acc;
}

void launch_interface_bprime(const RowsView& R, const double* val, const double* psi,
                             const double* src, double* bPrime, hipStream_t s)
{
    hipLaunchKernelGGL(k_interface_bprime, dim3(reduction_blocks(R.n)), dim3(FH_BLOCK), 0, s,
                       R, val, psi, src, bPrime);
}

// the interface entries' values: val[ifBase + e] = -coeffs[face[e]]
__global__ void k_interface_values(int32_t nE, const int32_t* __restrict__ face,
                                   const double* __restrict__ coeffs, double* __restrict__ tail)
{
    const int32_t e = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (e < nE) tail[e] = -coeffs[face[e]];
}

void launch_interface_values(int32_t nE, const int32_t* face, const double* coeffs,
                             double* tail, hipStream_t s)
{
    if (nE <= 0) return;
    hipLaunchKernelGGL(k_interface_values, dim3((nE + FH_BLOCK - 1)/FH_BLOCK), dim3(FH_BLOCK), 0,
                       s, nE, face, coeffs, tail);
}

void launch_sum_a(const RowsView& R, const double* diag, const double* val,
                  double* sumA, hipStream_t s)
{
    hipLaunchKernelGGL(k_sum_a, dim3(reduction_blocks(R.n)), dim3(FH_BLOCK), 0, s,
                       R, diag, val, sumA);
}

template<class T>
__global__ void k_reciprocal(int32_t n, T* a)
{
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) a[i] = T(1)/a[i];
}

template<class T>
void launch_reciprocal(int32_t n, T* a, hipStream_t s)
{
    hipLaunchKernelGGL(k_reciprocal<T>, dim3(reduction_blocks(n)), dim3(FH_BLOCK), 0, s, n, a);
}

// --------------------------------------------------------------------------
// Pipelined partitioned level sweep.
//
// One workgroup owns one part and walks its level segments in order
// (forward) or reverse order (backward).  Before a level it waits until
// every part it depends on has finished the levels its rows can read:
// forward, all levels <= L-1 of the parts holding lower neighbours;
// backward, all levels >= L+1 of the parts holding upper neighbours.
// Parts are handed out through a ticket counter in a topological order, so a
// waiting part only ever waits for parts that already hold a ticket, and
// every wait is bounded (on expiry the solve reports FH_ERR_SCHEDULE).
//
// Progress values carry the launch epoch: v = epoch*(nLevels+2) + k with
// k in [0, nLevels+1], so values of earlier launches never satisfy a wait.
// --------------------------------------------------------------------------

// one poll costs about a microsecond; a legitimate wait lasts at most the
// sweep itself (milliseconds), so 2^20 polls (about a second) means a
// broken schedule: the wait is abandoned and the solve reports it
#define FH_SPIN_LIMIT (1u << 20)

__device__ __forceinline__ unsigned long long ld_progress(const unsigned long long* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void st_progress(unsigned long long* p, unsigned long long v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// value loads / stores of the sweeps: plain inside one workgroup's fenced
// levels, device-coherent in the sync-free row sweep
template<bool COH, class T>
__device__ __forceinline__ T sw_ld(const T* p)
{
    if constexpr (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *p;
}

template<bool COH, class T>
__device__ __forceinline__ void sw_st(T* p, T v)
{
    if constexpr (COH) __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *p = v;
}

template<class T, int KIND, bool COH = false>
__device__ __forceinline__ void sweep_row(const RowsView& R, const SweepData<T>& d, int32_t r)
{
    const int32_t info = R.rowInfo[r];
    const int32_t nLo = info & 0xffff;
    const int32_t nEnt = info >> 16;
    const int32_t base = sell_base(R.sliceStart, r);
    if (KIND == SWEEP_PRECOND_FWD)
    {
        // DICPreconditioner.C:108-117 / DILUPreconditioner.C:117-128
        const T rDr = d.rD[r];
        T w = rDr*d.x[r];
        for (int32_t k = 0; k < nLo; k++)
        {
            const int32_t e = base + (k << 6);
            w -= rDr*d.val[e]*sw_ld<COH>(&d.w[R.col[e]]);
        }
        sw_st<COH>(&d.w[r], w);
    }
    else if (KIND == SWEEP_PRECOND_BWD)
    {
        // DICPreconditioner.C:119-122: faces in descending order
        const T rDr = d.rD[r];
        T w = d.w[r];
        for (int32_t k = nEnt - 1; k >= nLo; k--)
        {
            const int32_t e = base + (k << 6);
            w -= rDr*d.val[e]*sw_ld<COH>(&d.w[R.col[e]]);
        }
        sw_st<COH>(&d.w[r], w);
    }
    else if (KIND == SWEEP_PRECOND_BWD_T)
    {
        // DILUPreconditioner.C:175-184: faces visited in reverse losort
        // order, i.e. the row's upper neighbours by descending cell index
        const T rDr = d.rD[r];
        T w = d.w[r];
        int32_t lastNat = INT32_MAX;
        for (int32_t cnt = nLo; cnt < nEnt; cnt++)
        {
            int32_t best = -1, bestNat = -1;
            for (int32_t k = nLo; k < nEnt; k++)
            {
                const int32_t e = base + (k << 6);
                const int32_t nat = d.perm[R.col[e]];
                if (nat < lastNat && nat > bestNat) { bestNat = nat; best = e; }
            }
            lastNat = bestNat;
            w -= rDr*d.val[best]*sw_ld<COH>(&d.w[R.col[best]]);
        }
        sw_st<COH>(&d.w[r], w);
    }
    else if (KIND == SWEEP_RD_DIC)
    {
        // DICPreconditioner.C:67-72 (raw factor; reciprocal taken afterwards)
        T v = d.diag[r];
        for (int32_t k = 0; k < nLo; k++)
        {
            const int32_t e = base + (k << 6);
            const T u = d.val[e];
            v -= u*u/sw_ld<COH>(&d.w[R.col[e]]);
        }
        sw_st<COH>(&d.w[r], v);
    }
    else if (KIND == SWEEP_RD_DILU)
    {
        // DILUPreconditioner.C:67-72: upper[f]*lower[f]/rD[l]
        T v = d.diag[r];
        for (int32_t k = 0; k < nLo; k++)
        {
            const int32_t e = base + (k << 6);
            const T up = (T)d.upperF[R.face[e] >> 1];
            v -= up*d.val[e]/sw_ld<COH>(&d.w[R.col[e]]);
        }
        sw_st<COH>(&d.w[r], v);
    }
    else if (KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD)
    {
        // GaussSeidelSmoother.C:150-176 / symGaussSeidelSmoother.C:151-205:
        // bPrime accumulates the lower products in face order, then the
        // upper products are subtracted in face order, then / diag
        T t = d.x[r];
        for (int32_t k = 0; k < nEnt; k++)
        {
            const int32_t e = base + (k << 6);
            t -= d.val[e]*sw_ld<COH>(&d.w[R.col[e]]);
        }
        sw_st<COH>(&d.w[r], t/d.diag[r]);
    }
}

// One colour of a multi-colour Gauss-Seidel sweep: no row of a colour reads
// another row of the same colour, so the rows update in parallel with the
// row arithmetic of the ordered sweep (neighbours of earlier colours new,
// of later colours old)
template<class T>
__global__ __launch_bounds__(FH_BLOCK)
void k_sweep_colour(RowsView R, SweepData<T> d, const int32_t* __restrict__ rows, int32_t n,
                    const DevState* st)
{
    if (st && st->done) return;
    const int32_t i = blockIdx.x*FH_BLOCK + threadIdx.x;
    if (i < n) sweep_row<T, SWEEP_GS_FWD>(R, d, rows[i]);
}

template<class T>
void launch_sweep_colour(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                         int32_t n, const DevState* st, hipStream_t s)
{
    if (n <= 0) return;
    hipLaunchKernelGGL(k_sweep_colour<T>, dim3((n + FH_BLOCK - 1)/FH_BLOCK), dim3(FH_BLOCK), 0,
                       s, R, d, rows, n, st);
}

// All colours of a small mesh in one workgroup: nSweeps sweeps over the
// colours ascending (then descending when sym), a barrier between colours
template<class T>
__global__ __launch_bounds__(1024)
void k_sweep_colours_wg(RowsView R, SweepData<T> d, const int32_t* __restrict__ rows,
                        const int32_t* __restrict__ cstart, int32_t nCol, int32_t nSweeps,
                        int32_t sym, const DevState* st)
{
    if (st && st->done) return;
    const int32_t nPass = sym ? 2*nCol : nCol;
    for (int32_t s = 0; s < nSweeps; s++)
        for (int32_t q = 0; q < nPass; q++)
        {
            const int32_t k = q < nCol ? q : 2*nCol - 1 - q;
            const int32_t r0 = cstart[k], r1 = cstart[k + 1];
            for (int32_t i = r0 + (int32_t)threadIdx.x; i < r1; i += (int32_t)blockDim.x)
                sweep_row<T, SWEEP_GS_FWD>(R, d, rows[i]);
            __syncthreads();
        }
}

template<class T>
void launch_sweep_colours_wg(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                             const int32_t* cstart, int32_t nCol, int32_t nSweeps, bool sym,
                             const DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL(k_sweep_colours_wg<T>, dim3(1), dim3(1024), 0, s, R, d, rows, cstart,
                       nCol, nSweeps, (int32_t)sym, st);
}

// A run of consecutive small colours [k0, k1) in one workgroup, ascending
// or (desc) descending, a barrier between colours: one launch where each
// colour would have been its own
template<class T>
__global__ __launch_bounds__(1024)
void k_sweep_colour_run(RowsView R, SweepData<T> d, const int32_t* __restrict__ rows,
                        const int32_t* __restrict__ cstart, int32_t k0, int32_t k1,
                        int32_t desc, const DevState* st)
{
    if (st && st->done) return;
    for (int32_t q = 0; q < k1 - k0; q++)
    {
        const int32_t k = desc ? k1 - 1 - q : k0 + q;
        const int32_t r0 = cstart[k], r1 = cstart[k + 1];
        for (int32_t i = r0 + (int32_t)threadIdx.x; i < r1; i += (int32_t)blockDim.x)
            sweep_row<T, SWEEP_GS_FWD>(R, d, rows[i]);
        __syncthreads();
    }
}

template<class T>
void launch_sweep_colour_run(const RowsView& R, const SweepData<T>& d, const int32_t* rows,
                             const int32_t* cstart, int32_t k0, int32_t k1, bool desc,
                             const DevState* st, hipStream_t s)
{
    if (k1 <= k0) return;
    hipLaunchKernelGGL(k_sweep_colour_run<T>, dim3(1), dim3(1024), 0, s, R, d, rows, cstart, k0,
                       k1, (int32_t)desc, st);
}

template void launch_sweep_colour_run<double>(const RowsView&, const SweepData<double>&,
                                              const int32_t*, const int32_t*, int32_t, int32_t,
                                              bool, const DevState*, hipStream_t);
template void launch_sweep_colour_run<float>(const RowsView&, const SweepData<float>&,
                                             const int32_t*, const int32_t*, int32_t, int32_t,
                                             bool, const DevState*, hipStream_t);

template void launch_sweep_colours_wg<double>(const RowsView&, const SweepData<double>&,
                                              const int32_t*, const int32_t*, int32_t, int32_t,
                                              bool, const DevState*, hipStream_t);
template void launch_sweep_colours_wg<float>(const RowsView&, const SweepData<float>&,
                                             const int32_t*, const int32_t*, int32_t, int32_t,
                                             bool, const DevState*, hipStream_t);

template void launch_sweep_colour<double>(const RowsView&, const SweepData<double>&,
                                          const int32_t*, int32_t, const DevState*, hipStream_t);
template void launch_sweep_colour<float>(const RowsView&, const SweepData<float>&,
                                         const int32_t*, int32_t, const DevState*, hipStream_t);

template<class T, int KIND, bool BACKWARD>
__global__ __launch_bounds__(FH_BLOCK)
void k_sweep_parts(RowsView R, SchedView S, SweepSync sync, SweepData<T> d,
                   DevState* st)
{
    __shared__ int32_t s_part;
    __shared__ int32_t s_stop;
    if (threadIdx.x == 0)
    {
        const unsigned long long t =
            atomicAdd(sync.ticket, 1ull) - sync.epoch*(unsigned long long)S.nParts;
        if (t >= (unsigned long long)S.nParts)
        {
            // ticket counter out of step with the launch count: never index
            s_part = -1;
            if (st) st->scheduleError = 1;
        }
        else
        {
            const int32_t ti = (int32_t)t;
            s_part = BACKWARD ? S.order[S.nParts - 1 - ti] : S.order[ti];
        }
        s_stop = (st && st->done) ? 1 : 0;
    }
    __syncthreads();
    const int32_t p = s_part;
    if (p < 0) return;
    const unsigned long long K = (unsigned long long)S.nLevels + 2;
    const unsigned long long base = sync.epoch*K;
    unsigned long long* progress = sync.progress;
    const int32_t off = S.partLevelOff[p];
    const int32_t nSeg = S.partLevelOff[p + 1] - off - 1;
    const int32_t minL = S.partMinLevel[p];

    if (s_stop)
    {
        // still publish completion so that nothing waits on this part
        if (threadIdx.x == 0) st_progress(&progress[p], base + K - 1);
        return;
    }

    const int32_t depBeg = BACKWARD ? S.downOff[p] : S.upOff[p];
    const int32_t nDep = (BACKWARD ? S.downOff[p + 1] : S.upOff[p + 1]) - depBeg;
    const int32_t* depList = BACKWARD ? S.downList : S.upList;
    // thread t < nDep tracks dependency t; its last observed progress value
    unsigned long long seen = 0;
    int32_t myDep = -1;
    if ((int32_t)threadIdx.x < nDep) myDep = depList[depBeg + threadIdx.x];
    bool expired = false;
    const bool multi = S.nParts > 1;
    // levels below the first (above the last, backward) are trivially done
    if (multi && threadIdx.x == 0)
    {
        const int32_t maxL = minL + nSeg - 1;
        st_progress(&progress[p],
                    base + (BACKWARD ? (unsigned long long)(S.nLevels - maxL)
                                     : (unsigned long long)minL));
    }

    for (int32_t s = 0; s < nSeg; s++)
    {
        const int32_t si = BACKWARD ? nSeg - 1 - s : s;
        const int32_t L = minL + si;
        const int32_t r0 = S.segStart[off + si], r1 = S.segStart[off + si + 1];
        if (r0 == r1) continue;
        if (myDep >= 0)
        {
            // forward: need k >= L  (all levels <= L-1 done, k = level+1)
            // backward: need k >= nLevels - L  (all levels >= L+1 done)
            const unsigned long long need =
                base + (BACKWARD ? (unsigned long long)(S.nLevels - L)
                                 : (unsigned long long)L);
            uint32_t spins = 0;
            while (seen < need && !expired)
            {
                seen = ld_progress(&progress[myDep]);
                if (seen >= need) break;
                if (++spins > FH_SPIN_LIMIT) { expired = true; break; }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (nDep > 0)
        {
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        for (int32_t r = r0 + (int32_t)threadIdx.x; r < r1; r += FH_BLOCK)
            sweep_row<T, KIND>(R, d, r);
        if (multi) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (multi && threadIdx.x == 0)
            st_progress(&progress[p],
                        base + (BACKWARD ? (unsigned long long)(S.nLevels - L + 1)
                                         : (unsigned long long)(L + 1)));
    }
    if (threadIdx.x == 0) st_progress(&progress[p], base + K - 1);
    if (expired && st) st->scheduleError = 1;
}

// A row of the chunk sweep in two halves, for the Gauss-Seidel and the
// DIC/DILU application kinds.  Before the chunk's wait: the row's shape and
// scalars (source and diagonal, or rD and x / the forward result), the
// columns and coefficients of its first ROW_PREFETCH entries (32: with 16,
// the coarse GAMG levels' longer rows loaded their last entries one by one
// after the wait and set the pace, 0.200 against 0.163 s for three
// configuration-4 V-cycles), and (Gauss-
// Seidel) the values of those entries that are not dependencies of this
// sweep (upper neighbours forward, lower ones backward: no chunk of this
// launch writes them before this row is done, so they hold the previous
// values).  After the wait: the dependency values, all loads in flight
// together, then the row's arithmetic in the order of sweep_row
// (GaussSeidelSmoother.C:150-176, DICPreconditioner.C:108-122).
#define ROW_PREFETCH 32

template<int KIND>
struct RowPreKind
{
    static constexpr bool gs = KIND == SWEEP_GS_FWD || KIND == SWEEP_GS_BWD;
    static constexpr bool any = gs || KIND == SWEEP_PRECOND_FWD || KIND == SWEEP_PRECOND_BWD;
};

template<class T>
struct RowPre
{
    int32_t nLo = 0, nEnt = 0, base = 0;
    T a = T(0), b = T(1);       // GS: source, diagonal; DIC/DILU: x or w(r), rD
    int32_t col[ROW_PREFETCH];
    T val[ROW_PREFETCH], w[ROW_PREFETCH];
};

template<class T, int KIND, bool BACKWARD>
__device__ __forceinline__ void row_prefetch(const RowsView& R, const SweepData<T>& d,
                                             const T* wIn, int32_t r, bool have, RowPre<T>& p)
{
    if (have)
    {
        const int32_t info = R.rowInfo[r];
        p.nLo = info & 0xffff;
        p.nEnt = info >> 16;
        p.base = sell_base(R.sliceStart, r);
        if constexpr (RowPreKind<KIND>::gs)
        {
            p.a = d.x[r];
            p.b = d.diag[r];
        }
        else
        {
            // backward: this row's forward result, written by an earlier launch
            p.a = KIND == SWEEP_PRECOND_FWD ? d.x[r] : wIn[r];
            p.b = d.rD[r];
        }
    }
#pragma unroll
    for (int k = 0; k < ROW_PREFETCH; k++)
    {
        const bool in = k < p.nEnt;
        const int32_t e = p.base + (k << 6);
        p.col[k] = in ? R.col[e] : 0;
        p.val[k] = in ? d.val[e] : T(0);
    }
    if constexpr (RowPreKind<KIND>::gs)
    {
#pragma unroll
        for (int k = 0; k < ROW_PREFETCH; k++)
        {
            const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
            p.w[k] = k < p.nEnt && !dep ? wIn[p.col[k]] : T(0);
        }
    }
}

// the signalling-NaN sentinel of the value-polling passes (see
// k_sweep_chunks_vf)
template<class T> struct SweepSentinel;
template<> struct SweepSentinel<double>
{
    static constexpr long long bits = 0x7FF0DEAD5EED0001ll;
    __device__ static double value() { return __longlong_as_double(bits); }
    __device__ static bool is(double v) { return __double_as_longlong(v) == bits; }
};
template<> struct SweepSentinel<float>
{
    static constexpr int bits = 0x7FA05EED;
    __device__ static float value() { return __int_as_float(bits); }
    __device__ static bool is(float v) { return __float_as_int(v) == bits; }
};

// A row's entries past the prefetched ones, [kLo, kHi), 16 at a time with
// every load of a block in flight (columns and coefficients, then the
// values): ascending for Gauss-Seidel and the forward pass, descending for
// the backward pass.  Dependency values come from wDep (with VF: polled
// until no longer the sentinel), the others from wIn.  acc -= c*w
// (Gauss-Seidel) or acc -= rD*c*w.
template<class T, int KIND, bool BACKWARD, bool VF>
__device__ __forceinline__ T row_tail(const RowsView& R, const SweepData<T>& d, const T* wIn,
                                      const T* wDep, const RowPre<T>& p, int32_t kLo,
                                      int32_t kHi, T acc, T rDr, unsigned long long t0,
                                      unsigned long long waitTicks, bool& expired)
{
    constexpr int B = 16;
    constexpr bool desc = KIND == SWEEP_PRECOND_BWD;
    for (int32_t b = 0; b < kHi - kLo; b += B)
    {
        int32_t col[B];
        T val[B], w[B];
#pragma unroll
        for (int j = 0; j < B; j++)
        {
            const int32_t k = desc ? kHi - 1 - b - j : kLo + b + j;
            const bool in = desc ? k >= kLo : k < kHi;
            const int32_t e = p.base + (k << 6);
            col[j] = in ? R.col[e] : 0;
            val[j] = in ? d.val[e] : T(0);
        }
#pragma unroll
        for (int j = 0; j < B; j++)
        {
            const int32_t k = desc ? kHi - 1 - b - j : kLo + b + j;
            const bool in = desc ? k >= kLo : k < kHi;
            const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
            w[j] = !in ? T(0) : dep ? sw_ld<true>(&wDep[col[j]]) : wIn[col[j]];
        }
        if constexpr (VF)
        {
            for (;;)
            {
                bool pending = false;
#pragma unroll
                for (int j = 0; j < B; j++)
                {
                    const int32_t k = desc ? kHi - 1 - b - j : kLo + b + j;
                    const bool in = desc ? k >= kLo : k < kHi;
                    const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                    pending |= in && dep && SweepSentinel<T>::is(w[j]);
                }
                if (!pending || expired) break;
                if ((unsigned long long)(wall_clock64() - t0) > waitTicks)
                {
                    expired = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int j = 0; j < B; j++)
                {
                    const int32_t k = desc ? kHi - 1 - b - j : kLo + b + j;
                    const bool in = desc ? k >= kLo : k < kHi;
                    const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                    if (in && dep && SweepSentinel<T>::is(w[j])) w[j] = sw_ld<true>(&wDep[col[j]]);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < B; j++)
        {
            const int32_t k = desc ? kHi - 1 - b - j : kLo + b + j;
            const bool in = desc ? k >= kLo : k < kHi;
            if (in)
            {
                if constexpr (RowPreKind<KIND>::gs) acc -= val[j]*w[j];
                else acc -= rDr*val[j]*w[j];
            }
        }
    }
    return acc;
}

template<class T, int KIND, bool BACKWARD>
__device__ __forceinline__ void row_finish(const RowsView& R, const SweepData<T>& d, int32_t r,
                                           RowPre<T>& p)
{
#pragma unroll
    for (int k = 0; k < ROW_PREFETCH; k++)
    {
        const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
        if (k < p.nEnt && dep) p.w[k] = sw_ld<true>(&d.w[p.col[k]]);
    }
    if constexpr (RowPreKind<KIND>::gs)
    {
        T t = p.a;
#pragma unroll
        for (int k = 0; k < ROW_PREFETCH; k++)
            if (k < p.nEnt) t -= p.val[k]*p.w[k];
        bool unused = false;
        t = row_tail<T, KIND, BACKWARD, false>(R, d, d.w, d.w, p, ROW_PREFETCH, p.nEnt, t, T(0),
                                               0, 0, unused);
        sw_st<true>(&d.w[r], t/p.b);
    }
    else if constexpr (KIND == SWEEP_PRECOND_FWD)
    {
        // lower entries ascending
        const T rDr = p.b;
        T w = rDr*p.a;
#pragma unroll
        for (int k = 0; k < ROW_PREFETCH; k++)
            if (k < p.nLo) w -= rDr*p.val[k]*p.w[k];
        bool unused = false;
        w = row_tail<T, KIND, BACKWARD, false>(R, d, d.w, d.w, p, ROW_PREFETCH, p.nLo, w, rDr,
                                               0, 0, unused);
        sw_st<true>(&d.w[r], w);
    }
    else
    {
        // upper entries descending
        const T rDr = p.b;
        T w = p.a;
        // entries past the prefetched ones, down to the first upper entry
        const int32_t kEnd = p.nLo > ROW_PREFETCH ? p.nLo : ROW_PREFETCH;
        bool unused = false;
        w = row_tail<T, KIND, BACKWARD, false>(R, d, d.w, d.w, p, kEnd, p.nEnt, w, rDr, 0, 0,
                                               unused);
#pragma unroll
        for (int k = ROW_PREFETCH - 1; k >= 0; k--)
            if (k >= p.nLo && k < p.nEnt) w -= rDr*p.val[k]*p.w[k];
        sw_st<true>(&d.w[r], w);
    }
}

// --------------------------------------------------------------------------
// Sync-free chunk sweep (meshes without a tile or part structure, e.g. GAMG
// coarse levels).  The library numbering of such a mesh is sorted by
// dependency level; each level's rows are cut into chunks of at most 64,
// so a chunk's rows are independent and every row it reads in the sweep
// direction lies in an earlier chunk.  Persistent wavefronts take chunks
// through a ticket counter in sweep order, wait until the completion flags
// of the chunk's dependency chunks (precomputed per mesh) carry this
// launch's value (epoch + 1), compute one row per lane with sweep_row's
// arithmetic, and publish the chunk flag after the row stores completed.
// A wave only waits on chunks ticketed before its own, whose holders are
// running, so the sweep cannot deadlock; every wait is bounded in time
// and an expired wait is reported (and the chunk still published, so that
// the launch drains).
// --------------------------------------------------------------------------
template<class T, int KIND, bool BACKWARD>
__global__ __launch_bounds__(FH_BLOCK)
void k_sweep_chunks(RowsView R, SweepData<T> d, ChunkView C, RowSync sync,
                    unsigned int doneVal, DevState* st)
{
    // a finished solve still takes its tickets, so that the counter stays
    // in step with the host's ticketBase
    const bool stop = st && st->done;
    const int lane = (int)(threadIdx.x & 63);
    const int32_t* depOff = BACKWARD ? C.bwdOff : C.fwdOff;
    const int32_t* deps = BACKWARD ? C.bwdDeps : C.fwdDeps;
    for (;;)
    {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(sync.ticket, 1ull) - sync.ticketBase;
        t = __shfl(t, 0, 64);
        if (t >= (unsigned long long)C.nChunks) return;
        if (stop) continue;
        const int32_t c = BACKWARD ? C.nChunks - 1 - (int32_t)t : (int32_t)t;
        const int32_t q0 = depOff[c], q1 = depOff[c + 1];
        const int32_t r = C.start[c] + lane;
        const bool have = r < C.start[c + 1];
        RowPre<T> pre;
        if constexpr (RowPreKind<KIND>::any) row_prefetch<T, KIND, BACKWARD>(R, d, d.w, r, have, pre);
        bool expired = false;
        if (q1 > q0)
        {
            const unsigned long long t0 = wall_clock64();
            for (;;)
            {
                bool mine = true;
                for (int32_t q = q0 + lane; q < q1; q += 64)
                    if (__hip_atomic_load(&sync.flags[deps[q]], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT) < doneVal)
                    {
                        mine = false;
                        break;
                    }
                if (__all(mine)) break;
                if ((unsigned long long)(wall_clock64() - t0) > sync.waitTicks)
                {
                    expired = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        if (!expired && have)
        {
            if constexpr (RowPreKind<KIND>::any)
                row_finish<T, KIND, BACKWARD>(R, d, r, pre);
            else
                sweep_row<T, KIND, true>(R, d, r);
        }
        // the chunk's row stores complete before its flag is published
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (lane == 0)
        {
            if (expired)
            {
                __hip_atomic_store(sync.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (st) st->scheduleError = 1;
            }
            __hip_atomic_store(&sync.flags[c], doneVal, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// --------------------------------------------------------------------------
// Value-polling chunk sweep (the Gauss-Seidel and DIC/DILU kinds).  The
// pass writes into an output that holds a sentinel before the launch (a
// signalling NaN: no arithmetic produces one, a NaN result is quiet).  One
// lane of the wavefront polls the chunk's dependency written last (rows
// are numbered by level: the highest row forward, the lowest backward) in
// the output until it is no longer the sentinel; then every lane loads its
// dependency values and re-polls any still unwritten.  The hand-off
// between levels is the value itself: no completion flag, no wait for the
// stores.  Non-dependency values (and, backward, the row's own forward
// result) come from the input.  Chunks are still ticketed in sweep order,
// so a row only waits for rows of chunks whose holders are running; every
// wait is bounded in time, an expired wait is reported and the row still
// written (never the sentinel), so that the launch drains.
// --------------------------------------------------------------------------
// the fill and the copy back do nothing once the solve is done, like the
// pass itself (which then computes no row)
template<class T>
__global__ __launch_bounds__(FH_BLOCK)
void k_fill_sentinel(int32_t n, T* __restrict__ p, const DevState* st)
{
    const int32_t i = (int32_t)(blockIdx.x*FH_BLOCK + threadIdx.x);
    if (i < n && !(st && st->done)) p[i] = SweepSentinel<T>::value();
}

template<class T>
__global__ __launch_bounds__(FH_BLOCK)
void k_copy_rows(int32_t n, const T* __restrict__ a, T* __restrict__ b, const DevState* st)
{
    const int32_t i = (int32_t)(blockIdx.x*FH_BLOCK + threadIdx.x);
    if (i < n && !(st && st->done)) b[i] = a[i];
}

// one dependency value: poll until written (or the wait expires)
template<class T>
__device__ __forceinline__ T vf_poll(const T* a, unsigned long long t0,
                                     unsigned long long waitTicks, bool& expired)
{
    for (;;)
    {
        const T v = sw_ld<true>(a);
        if (!SweepSentinel<T>::is(v) || expired) return v;
        if ((unsigned long long)(wall_clock64() - t0) > waitTicks)
        {
            expired = true;
            return v;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

template<class T, int KIND, bool BACKWARD>
__global__ __launch_bounds__(FH_BLOCK)
void k_sweep_chunks_vf(RowsView R, SweepData<T> d, const T* wIn, T* wOut, ChunkView C,
                       RowSync sync, DevState* st)
{
    const bool stop = st && st->done;
    const int lane = (int)(threadIdx.x & 63);
    for (;;)
    {
        unsigned long long t = 0;
        if (lane == 0) t = atomicAdd(sync.ticket, 1ull) - sync.ticketBase;
        t = __shfl(t, 0, 64);
        if (t >= (unsigned long long)C.nChunks) return;
        if (stop) continue;
        const int32_t c = BACKWARD ? C.nChunks - 1 - (int32_t)t : (int32_t)t;
        const int32_t r = C.start[c] + lane;
        const bool have = r < C.start[c + 1];
        RowPre<T> p;
        row_prefetch<T, KIND, BACKWARD>(R, d, wIn, r, have, p);
        bool expired = false;
        const unsigned long long t0 = wall_clock64();
        // the chunk's dependency written last, polled by one lane
        {
            int32_t key = INT32_MIN;
#pragma unroll
            for (int k = 0; k < ROW_PREFETCH; k++)
            {
                const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                if (k < p.nEnt && dep)
                    key = max(key, BACKWARD ? -p.col[k] : p.col[k]);
            }
            for (int off = 32; off > 0; off >>= 1) key = max(key, __shfl_xor(key, off, 64));
            if (lane == 0 && key != INT32_MIN)
                (void)vf_poll(&wOut[BACKWARD ? -key : key], t0, sync.waitTicks, expired);
        }
        if (have)
        {
#pragma unroll
            for (int k = 0; k < ROW_PREFETCH; k++)
            {
                const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                if (k < p.nEnt && dep) p.w[k] = SweepSentinel<T>::value();
            }
            // the prefetched entries' dependencies: all loads in flight, then
            // again for those not yet written
            for (;;)
            {
#pragma unroll
                for (int k = 0; k < ROW_PREFETCH; k++)
                {
                    const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                    if (k < p.nEnt && dep && SweepSentinel<T>::is(p.w[k]))
                        p.w[k] = sw_ld<true>(&wOut[p.col[k]]);
                }
                bool pending = false;
#pragma unroll
                for (int k = 0; k < ROW_PREFETCH; k++)
                {
                    const bool dep = BACKWARD ? k >= p.nLo : k < p.nLo;
                    pending |= k < p.nEnt && dep && SweepSentinel<T>::is(p.w[k]);
                }
                if (!pending || expired) break;
                if ((unsigned long long)(wall_clock64() - t0) > sync.waitTicks)
                {
                    expired = true;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            T v;
            if constexpr (RowPreKind<KIND>::gs)
            {
                // GaussSeidelSmoother.C:150-176: entries in order
                T tt = p.a;
#pragma unroll
                for (int k = 0; k < ROW_PREFETCH; k++)
                    if (k < p.nEnt) tt -= p.val[k]*p.w[k];
                tt = row_tail<T, KIND, BACKWARD, true>(R, d, wIn, wOut, p, ROW_PREFETCH, p.nEnt,
                                                       tt, T(0), t0, sync.waitTicks, expired);
                v = tt/p.b;
            }
            else if constexpr (KIND == SWEEP_PRECOND_FWD)
            {
                // DICPreconditioner.C:108-117: lower entries ascending
                const T rDr = p.b;
                T w = rDr*p.a;
#pragma unroll
                for (int k = 0; k < ROW_PREFETCH; k++)
                    if (k < p.nLo) w -= rDr*p.val[k]*p.w[k];
                w = row_tail<T, KIND, BACKWARD, true>(R, d, wIn, wOut, p, ROW_PREFETCH, p.nLo, w,
                                                      rDr, t0, sync.waitTicks, expired);
                v = w;
            }
            else
            {
                // DICPreconditioner.C:119-122: upper entries descending
                const T rDr = p.b;
                T w = p.a;
                const int32_t kEnd = p.nLo > ROW_PREFETCH ? p.nLo : ROW_PREFETCH;
                w = row_tail<T, KIND, BACKWARD, true>(R, d, wIn, wOut, p, kEnd, p.nEnt, w, rDr,
                                                      t0, sync.waitTicks, expired);
#pragma unroll
                for (int k = ROW_PREFETCH - 1; k >= 0; k--)
                    if (k >= p.nLo && k < p.nEnt) w -= rDr*p.val[k]*p.w[k];
                v = w;
            }
            if (expired)
            {
                // never the sentinel: later rows drain
                if (SweepSentinel<T>::is(v)) v = T(0);
                __hip_atomic_store(sync.err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (st) st->scheduleError = 1;
            }
            sw_st<true>(&wOut[r], v);
        }
    }
}

template<class T, int KIND, bool BACKWARD>
static void launch_rows_vf(const RowsView& R, const SweepData<T>& d, const T* wIn, T* wOut,
                           int32_t nRows, const ChunkView& C, RowSync& sync, DevState* st,
                           hipStream_t s)
{
    const unsigned int g = (unsigned int)((nRows + FH_BLOCK - 1)/FH_BLOCK);
    if (nRows > 0)
        hipLaunchKernelGGL(k_fill_sentinel<T>, dim3(g), dim3(FH_BLOCK), 0, s, nRows, wOut,
                           (const DevState*)st);
    const int32_t wantBlocks = (C.nChunks + 3)/4;
    const int32_t grid = wantBlocks < kRowSweepBlocks ? wantBlocks : kRowSweepBlocks;
    hipLaunchKernelGGL((k_sweep_chunks_vf<T, KIND, BACKWARD>), dim3(grid), dim3(FH_BLOCK), 0, s,
                       R, d, wIn, wOut, C, sync, st);
    sync.ticketBase += (unsigned long long)C.nChunks + 4ull*(unsigned long long)grid;
    sync.epoch++;
}

// in-place passes through the scratch rows: sentinel-filled output, then
// copied back
template<class T, int KIND, bool BACKWARD>
static void launch_rows_vf_inplace(const RowsView& R, const SweepData<T>& d, int32_t nRows,
                                   T* tmp, const ChunkView& C, RowSync& sync, DevState* st,
                                   hipStream_t s)
{
    launch_rows_vf<T, KIND, BACKWARD>(R, d, d.w, tmp, nRows, C, sync, st, s);
    const unsigned int g = (unsigned int)((nRows + FH_BLOCK - 1)/FH_BLOCK);
    if (nRows > 0)
        hipLaunchKernelGGL(k_copy_rows<T>, dim3(g), dim3(FH_BLOCK), 0, s, nRows,
                           (const T*)tmp, d.w, (const DevState*)st);
}

template<class T, int KIND, bool BACKWARD>
static void launch_rows_kind(const RowsView& R, const SweepData<T>& d, const ChunkView& C,
                             RowSync& sync, DevState* st, hipStream_t s)
{
    const int32_t wantBlocks = (C.nChunks + 3)/4;
    const int32_t grid = wantBlocks < kRowSweepBlocks ? wantBlocks : kRowSweepBlocks;
    const unsigned int doneVal = (unsigned int)(sync.epoch + 1);
    hipLaunchKernelGGL((k_sweep_chunks<T, KIND, BACKWARD>), dim3(grid), dim3(FH_BLOCK), 0, s,
                       R, d, C, sync, doneVal, st);
    // every wavefront takes one ticket past the last chunk
    sync.ticketBase += (unsigned long long)C.nChunks + 4ull*(unsigned long long)grid;
    sync.epoch++;
}

template<class T>
void launch_sweep_rows(int kind, const RowsView& R, const SweepData<T>& d, const ChunkView& C,
                       RowSync& sync, DevState* st, hipStream_t s, int32_t nRows, T* tmp)
{
    if (tmp)
    {
        switch (kind)
        {
            case SWEEP_GS_FWD:
                launch_rows_vf_inplace<T, SWEEP_GS_FWD, false>(R, d, nRows, tmp, C, sync, st, s);
                return;
            case SWEEP_GS_BWD:
                launch_rows_vf_inplace<T, SWEEP_GS_BWD, true>(R, d, nRows, tmp, C, sync, st, s);
                return;
            case SWEEP_PRECOND_FWD:
                // w is output only, unless it is also the input
                if (d.w != d.x)
                    launch_rows_vf<T, SWEEP_PRECOND_FWD, false>(R, d, d.w, d.w, nRows, C, sync,
                                                                st, s);
                else
                    launch_rows_vf_inplace<T, SWEEP_PRECOND_FWD, false>(R, d, nRows, tmp, C,
                                                                        sync, st, s);
                return;
            case SWEEP_PRECOND_BWD:
                launch_rows_vf_inplace<T, SWEEP_PRECOND_BWD, true>(R, d, nRows, tmp, C, sync,
                                                                   st, s);
                return;
            default:
                break;
        }
    }
    switch (kind)
    {
        case SWEEP_PRECOND_FWD:   launch_rows_kind<T, SWEEP_PRECOND_FWD, false>(R, d, C, sync, st, s); break;
        case SWEEP_PRECOND_BWD:   launch_rows_kind<T, SWEEP_PRECOND_BWD, true>(R, d, C, sync, st, s); break;
        case SWEEP_PRECOND_BWD_T: launch_rows_kind<T, SWEEP_PRECOND_BWD_T, true>(R, d, C, sync, st, s); break;
        case SWEEP_RD_DIC:        launch_rows_kind<T, SWEEP_RD_DIC, false>(R, d, C, sync, st, s); break;
        case SWEEP_RD_DILU:       launch_rows_kind<T, SWEEP_RD_DILU, false>(R, d, C, sync, st, s); break;
        case SWEEP_GS_FWD:        launch_rows_kind<T, SWEEP_GS_FWD, false>(R, d, C, sync, st, s); break;
        case SWEEP_GS_BWD:        launch_rows_kind<T, SWEEP_GS_BWD, true>(R, d, C, sync, st, s); break;
        default: fail(FH_ERR_INVALID, "unknown sweep kind %d", kind);
    }
}

template void launch_sweep_rows<double>(int, const RowsView&, const SweepData<double>&, const ChunkView&, RowSync&, DevState*, hipStream_t, int32_t, double*);
template void launch_sweep_rows<float>(int, const RowsView&, const SweepData<float>&, const ChunkView&, RowSync&, DevState*, hipStream_t, int32_t, float*);

template<class T, int KIND, bool BACKWARD>
static void launch_sweep_kind(const RowsView& R, const SchedView& S, SweepSync& sync,
                              const SweepData<T>& d, DevState* st, hipStream_t s)
{
    hipLaunchKernelGGL((k_sweep_parts<T, KIND, BACKWARD>), dim3(S.nParts), dim3(FH_BLOCK),
                       0, s, R, S, sync, d, st);
    sync.epoch++;
}

template<class T>
void launch_sweep(int kind, const RowsView& R, const SchedView& S, SweepSync& sync,
                  const SweepData<T>& d, DevState* st, hipStream_t s)
{
    switch (kind)
    {
        case SWEEP_PRECOND_FWD: launch_sweep_kind<T, SWEEP_PRECOND_FWD, false>(R, S, sync, d, st, s); break;
        case SWEEP_PRECOND_BWD: launch_sweep_kind<T, SWEEP_PRECOND_BWD, true>(R, S, sync, d, st, s); break;
        case SWEEP_RD_DIC:      launch_sweep_kind<T, SWEEP_RD_DIC, false>(R, S, sync, d, st, s); break;
        case SWEEP_RD_DILU:     launch_sweep_kind<T, SWEEP_RD_DILU, false>(R, S, sync, d, st, s); break;
        case SWEEP_GS_FWD:      launch_sweep_kind<T, SWEEP_GS_FWD, false>(R, S, sync, d, st, s); break;
        case SWEEP_GS_BWD:      launch_sweep_kind<T, SWEEP_GS_BWD, true>(R, S, sync, d, st, s); break;
        case SWEEP_PRECOND_BWD_T: launch_sweep_kind<T, SWEEP_PRECOND_BWD_T, true>(R, S, sync, d, st, s); break;
        default: fail(FH_ERR_INVALID, "unknown sweep kind %d", kind);
    }
}

// ---- streaming probes (achievable bandwidth) -----------------------------
__global__ void k_stream_read(const float4* __restrict__ a, int64_t n4, float* out)
{
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t i = blockIdx.x*(int64_t)blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x*blockDim.x)
    {
        const float4 v = a[i];
        s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    // never true for the zero-filled input: keeps the loads alive
    if (s.x + s.y + s.z + s.w == 1.0f) out[0] = s.x;
}

__global__ void k_stream_copy(const float4* __restrict__ a, float4* __restrict__ b, int64_t n4)
{
    for (int64_t i = blockIdx.x*(int64_t)blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x*blockDim.x)
        b[i] = a[i];
}

__global__ void k_stream_3r1w(const float4* __restrict__ a, const float4* __restrict__ c,
                              const float4* __restrict__ d, float4* __restrict__ b, int64_t n4)
{
    for (int64_t i = blockIdx.x*(int64_t)blockDim.x + threadIdx.x; i < n4;
         i += (int64_t)gridDim.x*blockDim.x)
    {
        const float4 x = a[i], y = c[i], z = d[i];
        b[i] = make_float4(x.x + y.x + z.x, x.y + y.y + z.y, x.z + y.z + z.z, x.w + y.w + z.w);
    }
}

void launch_stream_probe(int pattern, int64_t n4, int grid, float4* const buf[4], float* out,
                         hipStream_t s)
{
    if (pattern == 0)
        hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(FH_BLOCK), 0, s, buf[0], n4, out);
    else if (pattern == 1)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(FH_BLOCK), 0, s, buf[0], buf[1], n4);
    else
        hipLaunchKernelGGL(k_stream_3r1w, dim3(grid), dim3(FH_BLOCK), 0, s, buf[0], buf[2],
                           buf[3], buf[1], n4);
}

// explicit instantiations
template void launch_gather_values<double>(const RowsView&, int64_t, const double*, const double*, double*, hipStream_t);
template void launch_gather_values<float>(const RowsView&, int64_t, const double*, const double*, float*, hipStream_t);
template void launch_permute_in<double, double>(int32_t, const int32_t*, const double*, double*, hipStream_t);
template void launch_permute_out<double, double>(int32_t, const int32_t*, const double*, double*, hipStream_t);
template void launch_permute_in<float, float>(int32_t, const int32_t*, const float*, float*, hipStream_t);
template void launch_permute_out<float, float>(int32_t, const int32_t*, const float*, float*, hipStream_t);
template void launch_convert<double, float>(int64_t, const double*, float*, hipStream_t);
template void launch_convert<float, double>(int64_t, const float*, double*, hipStream_t);
template void launch_amul<double>(const RowsView&, const double*, const double*, const double*, double*, const DevState*, hipStream_t);
template void launch_amul<float>(const RowsView&, const float*, const float*, const float*, float*, const DevState*, hipStream_t);
template void launch_tmul<double>(const RowsView&, const double*, const double*, const double*, double*, hipStream_t);
template void launch_residual<double>(const RowsView&, const double*, const double*, const double*, const double*, double*, const DevState*, hipStream_t);
template void launch_residual<float>(const RowsView&, const float*, const float*, const float*, const float*, float*, const DevState*, hipStream_t);
template void launch_sweep<double>(int, const RowsView&, const SchedView&, SweepSync&, const SweepData<double>&, DevState*, hipStream_t);
template void launch_sweep<float>(int, const RowsView&, const SchedView&, SweepSync&, const SweepData<float>&, DevState*, hipStream_t);
template void launch_reciprocal<double>(int32_t, double*, hipStream_t);
template void launch_reciprocal<float>(int32_t, float*, hipStream_t);

} // namespace fh
